// Solver kernels for gfx950: node-block (3x3) SpMV and Jacobi-PCG, replacing
// the UMFPACK solve of solve_system! (src/solver.jl:42-43), plus the vector
// kernels of the device-resident stage driver.
//
// Vectors live in node order (3 dofs per node).  The operator is K with the
// prescribed dofs masked out: y = mask .* (K x).  With x zero at prescribed
// dofs this is K11 x on the unknowns, so one resident copy of K serves K11,
// K12 U2 (right-hand side) and K21 U1 + K22 U2 (reactions).  Block values
// are stored slot-major (value s of block k at s*nblk + k) so that the lanes
// of a wave, which walk consecutive blocks of a row, read contiguous memory.
//
// Reductions need no atomics and no fences: a reducing kernel writes one
// partial per workgroup, and every workgroup of the next kernel that needs
// the scalar sums those partials itself in index order (kernel boundaries
// make them visible).  Results are deterministic for a given model.
//
// Precision.  Plain path: FP64 PCG.  Mixed path: the same CG recurrences
// with FP64 vectors, FP64 dot products and FP64 solution update, but the
// SpMV reads FP32 values and an FP32 copy of the search direction and
// accumulates in FP32, and the Jacobi preconditioner is applied in FP32.
// The FP64 residual r = b - K x is recomputed with the FP64 matrix whenever
// the recursive residual has dropped 10x since the last recomputation and
// at the end (residual replacement), which keeps the iteration converging
// to the FP64 solution without restarting the Krylov process; a model whose
// FP32 operator drifts too far from K falls back to the FP64 operator.
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "amr_internal.h"

namespace {

constexpr int WGV = 256;  // vector kernels
constexpr int WGS = 512;  // SpMV kernels
constexpr int MAX_GR = 4096, MAX_GV = 512;

// scalar and control slots
enum { S_RZ = 0, S_RR, S_BB, S_THR, S_KX, S_RES, S_ERRN, S_ERRD, S_GAP, S_NSCAL = 16 };
// C_DONE is read by every workgroup at kernel entry and written only where
// all workgroups of the running launch agree on it (breakdown); the
// convergence test writes C_STOP, which the next SpMV promotes to C_DONE, so
// no workgroup ever sees a flag raised during its own launch.
enum { C_DONE = 0, C_ITERS, C_BREAK, C_NAN, C_STOP, C_NCTL = 16 };
// partial buffers inside m->part
enum { P_PQ = 0, P_RZ0 = 4096, P_RZ1 = 4608, P_RR0 = 5120, P_RR1 = 5632, P_A = 8192, P_B = 12288, P_C = 16384,
       P_D = 20480, P_END = 24576 };

template <int NT>
__device__ inline double block_sum(double v)
{
    __shared__ double sh[NT / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < NT / 64; i++) t += sh[i];
    return t;  // same value in every thread
}

template <int NT>
__device__ inline double block_max(double v)
{
    __shared__ double sh[NT / 64];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    __syncthreads();
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < NT / 64; i++) t = fmax(t, sh[i]);
    return t;
}

// Sum of n partials in a fixed order; every thread gets the result.
template <int NT>
__device__ inline double sum_partials(const double *__restrict__ P, int n)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) v += P[i];
    return block_sum<NT>(v);
}

// Preconditioner: Jacobi (inverse diagonal) or 3x3 nodal block Jacobi
// (inverse of each node's diagonal block restricted to its free dofs), in
// FP64 or applied in FP32 (mixed paths).
struct Prec {
    const double *dinv;
    const float *dinv32;
    const double *binv;   // 9 per node, node-major
    const float *binv32;
    int f32, block;
};

__device__ inline void apply_prec(const Prec &P, int64_t node, const double r[3], double z[3])
{
    const int64_t i = node * 3;
    if (P.block) {
        if (P.f32) {
            const float *B = P.binv32 + node * 9;
            const float r0 = (float)r[0], r1 = (float)r[1], r2 = (float)r[2];
            for (int d = 0; d < 3; d++) z[d] = (double)fmaf(B[d * 3 + 2], r2, fmaf(B[d * 3 + 1], r1, B[d * 3] * r0));
        } else {
            const double *B = P.binv + node * 9;
            for (int d = 0; d < 3; d++) z[d] = fma(B[d * 3 + 2], r[2], fma(B[d * 3 + 1], r[1], B[d * 3] * r[0]));
        }
    } else if (P.f32) {
        for (int d = 0; d < 3; d++) z[d] = (double)(P.dinv32[i + d] * (float)r[d]);
    } else {
        for (int d = 0; d < 3; d++) z[d] = P.dinv[i + d] * r[d];
    }
}

// ------------------------------------------------------------------ SpMV
//
// Row tiles: workgroup b owns block rows [b*R, (b+1)*R) (R a power of two
// fixed per model), stages their row pointers in LDS and walks them with L
// lanes per row.  Each row's sum is formed over 64 virtual slots (slot j
// takes blocks rb+j, rb+j+64, ... in order) reduced by one fixed binary
// tree, whatever L is: a lane holds 64/L slots and performs the tree levels
// above L itself, the levels below L are cross-lane shuffles.  Per-row dot
// contributions go to LDS and are reduced by a fixed tree too.  So neither
// the row sums nor the partials depend on L, and L can be autotuned.

template <int L, bool ABS, typename VT, typename XT, typename AT>
__device__ inline void row_sum(int32_t rb, int32_t re, int sub, const int32_t *__restrict__ col,
                               const VT *__restrict__ vals, int64_t nblk, const XT *__restrict__ x, AT y[3])
{
    constexpr int NS = 64 / L;
    AT acc[NS][3];
#pragma unroll
    for (int t = 0; t < NS; t++) acc[t][0] = acc[t][1] = acc[t][2] = (AT)0;
    for (int32_t base = rb; base < re; base += 64) {
#pragma unroll
        for (int t = 0; t < NS; t++) {
            const int32_t k = base + sub + L * t;
            if (k < re) {
                const int64_t c = __builtin_nontemporal_load(col + k);
                const AT x0 = (AT)x[c * 3], x1 = (AT)x[c * 3 + 1], x2 = (AT)x[c * 3 + 2];
                const VT *vp = vals + kv_at(k, 0);
                AT v[9];
#pragma unroll
                for (int q = 0; q < 9; q++) v[q] = (AT)__builtin_nontemporal_load(vp + q * 64);
                if (ABS) {
                    acc[t][0] += fabs(v[0] * x0) + fabs(v[1] * x1) + fabs(v[2] * x2);
                    acc[t][1] += fabs(v[3] * x0) + fabs(v[4] * x1) + fabs(v[5] * x2);
                    acc[t][2] += fabs(v[6] * x0) + fabs(v[7] * x1) + fabs(v[8] * x2);
                } else {
                    acc[t][0] = fma(v[2], x2, fma(v[1], x1, fma(v[0], x0, acc[t][0])));
                    acc[t][1] = fma(v[5], x2, fma(v[4], x1, fma(v[3], x0, acc[t][1])));
                    acc[t][2] = fma(v[8], x2, fma(v[7], x1, fma(v[6], x0, acc[t][2])));
                }
            }
        }
    }
    // tree levels with offset >= L combine slots held by the same lane
#pragma unroll
    for (int off = 32; off >= L; off >>= 1) {
#pragma unroll
        for (int t = 0; t < off / L; t++)
#pragma unroll
            for (int d = 0; d < 3; d++) acc[t][d] = acc[t][d] + acc[t + off / L][d];
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        AT v = acc[0][d];
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, L);
        y[d] = v;
    }
}

// Fixed-order tree sum of R (power of two) LDS values; result in every thread.
__device__ inline double lds_tree_sum(double *s, int R)
{
    for (int h = R / 2; h > 0; h >>= 1) {
        for (int i = threadIdx.x; i < h; i += blockDim.x) s[i] = s[i] + s[i + h];
        __syncthreads();
    }
    const double t = s[0];
    __syncthreads();
    return t;
}

// Dynamic LDS layout of the row-tile kernels: R+1 row pointers, then up to
// four R-long double arrays.
__device__ inline int32_t *tile_rowptr(char *sm) { return (int32_t *)sm; }
__device__ inline double *tile_dbl(char *sm, int R, int j)
{
    const int off = ((R + 1) * 4 + 15) & ~15;
    return (double *)(sm + off) + (int64_t)j * R;
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8).  Tile t(b) gives each XCD one contiguous run of row tiles, so
// the slice of x its rows gather stays in that XCD's L2.
__device__ inline int64_t xcd_tile(int64_t b, int64_t G)
{
    const int64_t x = b & 7, i = b >> 3, q = G >> 3, rm = G & 7;
    return x * q + (x < rm ? x : rm) + i;
}

__device__ inline int load_tile_rows(const int32_t *__restrict__ rowptr, int64_t nrows, int R, int32_t *srp,
                                     int64_t &r0)
{
    r0 = xcd_tile(blockIdx.x, gridDim.x) * R;
    const int nr = (int)std::min<int64_t>(R, nrows - r0);
    for (int i = threadIdx.x; i <= nr; i += blockDim.x) srp[i] = rowptr[r0 + i];
    __syncthreads();
    return nr;
}

// y = K x (mode 0) or mask .* (K x) (mode 1), FP64
template <int L>
__global__ void __launch_bounds__(WGS) spmv_kernel(int64_t nrows, int R, const int32_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ col, const double *__restrict__ vals,
                                                   int64_t nblk, const double *__restrict__ x, double *__restrict__ y,
                                                   const double *__restrict__ mask, int mode)
{
    extern __shared__ __align__(16) char sm[];
    int32_t *srp = tile_rowptr(sm);
    int64_t r0;
    const int nr = load_tile_rows(rowptr, nrows, R, srp, r0);
    constexpr int RPW = WGS / L;
    const int sub = threadIdx.x % L;
    for (int i = threadIdx.x / L; i < nr; i += RPW) {
        double yy[3];
        row_sum<L, false, double, double, double>(srp[i], srp[i + 1], sub, col, vals, nblk, x, yy);
        if (sub == 0) {
            const int64_t row = r0 + i;
            for (int d = 0; d < 3; d++)
                y[row * 3 + d] = (mode == 1 && mask[row * 3 + d] == 0.0) ? 0.0 : yy[d];
        }
    }
}

// r = mask .* (b - K x) in FP64; per-tile partials ||r||^2, ||b||^2,
// || |K||x| + |b| ||^2 and, with rold, the drift ||r - rold||^2
template <int L>
__global__ void __launch_bounds__(WGS) residual_kernel(int64_t nrows, int R, const int32_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ col,
                                                       const double *__restrict__ vals, int64_t nblk,
                                                       const double *__restrict__ x, const double *__restrict__ b,
                                                       double *__restrict__ r, const double *__restrict__ rold,
                                                       const double *__restrict__ mask, double *__restrict__ part)
{
    extern __shared__ __align__(16) char sm[];
    int32_t *srp = tile_rowptr(sm);
    double *s_rr = tile_dbl(sm, R, 0), *s_bb = tile_dbl(sm, R, 1), *s_kx = tile_dbl(sm, R, 2),
           *s_gap = tile_dbl(sm, R, 3);
    for (int i = threadIdx.x; i < R; i += WGS) s_rr[i] = s_bb[i] = s_kx[i] = s_gap[i] = 0.0;
    int64_t r0;
    const int nr = load_tile_rows(rowptr, nrows, R, srp, r0);
    constexpr int RPW = WGS / L;
    const int sub = threadIdx.x % L;
    for (int i = threadIdx.x / L; i < nr; i += RPW) {
        double y[3], a[3];
        row_sum<L, false, double, double, double>(srp[i], srp[i + 1], sub, col, vals, nblk, x, y);
        row_sum<L, true, double, double, double>(srp[i], srp[i + 1], sub, col, vals, nblk, x, a);
        if (sub == 0) {
            double rr = 0.0, bb = 0.0, kx = 0.0, gap = 0.0;
            for (int d = 0; d < 3; d++) {
                const int64_t k = (r0 + i) * 3 + d;
                const bool fr = mask[k] != 0.0;
                const double ri = fr ? b[k] - y[d] : 0.0;
                const double bi = fr ? b[k] : 0.0;
                const double ai = fr ? a[d] + fabs(b[k]) : 0.0;
                r[k] = ri;
                rr = fma(ri, ri, rr);
                bb = fma(bi, bi, bb);
                kx = fma(ai, ai, kx);
                if (rold) {
                    const double g = ri - rold[k];
                    gap = fma(g, g, gap);
                }
            }
            s_rr[i] = rr;
            s_bb[i] = bb;
            s_kx[i] = kx;
            s_gap[i] = gap;
        }
    }
    __syncthreads();
    const double trr = lds_tree_sum(s_rr, R), tbb = lds_tree_sum(s_bb, R), tkx = lds_tree_sum(s_kx, R),
                 tgap = lds_tree_sum(s_gap, R);
    if (threadIdx.x == 0) {
        const int64_t t = xcd_tile(blockIdx.x, gridDim.x);
        part[P_A + t] = trr;
        part[P_B + t] = tbb;
        part[P_C + t] = tkx;
        part[P_D + t] = tgap;
    }
}

// ------------------------------------------------------------------- PCG
//
// Iteration k (parity pi = k % 2), three kernels:
//   pcg_spmv:   q = A p, partial p.q                          -> P_PQ
//   pcg_update: a = rz_k / pq; x += a p; r -= a q; z = M r;
//               partials rz_{k+1}, rr                         -> P_RZ[pi], P_RR[pi]
//   pcg_dir:    b = rz_{k+1} / rz_k; p = z + b p; convergence test
// rz_k is the sum of P_RZ[1-pi] (written by iteration k-1, by the
// initialisation, or by a residual replacement).

template <int L, typename VT, typename XT, typename AT>
__global__ void __launch_bounds__(WGS) pcg_spmv_kernel(int64_t nrows, int R, const int32_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ col,
                                                       const VT *__restrict__ vals, int64_t nblk,
                                                       const XT *__restrict__ px, const double *__restrict__ p,
                                                       double *__restrict__ q, const double *__restrict__ mask,
                                                       double *__restrict__ part, int32_t *__restrict__ ctl)
{
    if (ctl[C_DONE]) return;
    if (ctl[C_STOP]) {
        // converged in the previous iteration: every workgroup leaves, and
        // the first one raises C_DONE for the kernels that follow
        if (blockIdx.x == 0 && threadIdx.x == 0) ctl[C_DONE] = 1;
        return;
    }
    extern __shared__ __align__(16) char sm[];
    int32_t *srp = tile_rowptr(sm);
    double *sdot = tile_dbl(sm, R, 0);
    static_assert(sizeof(AT) >= sizeof(VT), "accumulate at least in the storage precision");
    for (int i = threadIdx.x; i < R; i += WGS) sdot[i] = 0.0;
    int64_t r0;
    const int nr = load_tile_rows(rowptr, nrows, R, srp, r0);
    constexpr int RPW = WGS / L;
    const int sub = threadIdx.x % L;
    for (int i = threadIdx.x / L; i < nr; i += RPW) {
        AT y[3];
        row_sum<L, false, VT, XT, AT>(srp[i], srp[i + 1], sub, col, vals, nblk, px, y);
        if (sub == 0) {
            double c = 0.0;
            for (int d = 0; d < 3; d++) {
                const int64_t k = (r0 + i) * 3 + d;
                const double qk = mask[k] != 0.0 ? (double)y[d] : 0.0;
                q[k] = qk;
                c = fma(p[k], qk, c);
            }
            sdot[i] = c;
        }
    }
    __syncthreads();
    const double pq = lds_tree_sum(sdot, R);
    if (threadIdx.x == 0) part[P_PQ + xcd_tile(blockIdx.x, gridDim.x)] = pq;
}

__global__ void __launch_bounds__(WGV) pcg_update_kernel(int64_t nnodes, double *__restrict__ x,
                                                         double *__restrict__ r, double *__restrict__ z,
                                                         const double *__restrict__ p, const double *__restrict__ q,
                                                         const Prec P, double *__restrict__ part, int npq, int nv,
                                                         int pi, int32_t *__restrict__ ctl)
{
    if (ctl[C_DONE]) return;
    const double pq = sum_partials<WGV>(part + P_PQ, npq);
    const double rz = sum_partials<WGV>(part + (pi ? P_RZ0 : P_RZ1), nv);
    if (!(pq > 0.0) || !std::isfinite(pq)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            ctl[C_BREAK] = 1;
            ctl[C_DONE] = 1;
        }
        return;
    }
    const double a = rz / pq;
    double rzn = 0.0, rr = 0.0;
    for (int64_t nd = (int64_t)blockIdx.x * WGV + threadIdx.x; nd < nnodes; nd += (int64_t)gridDim.x * WGV) {
        double rv[3], zv[3];
        for (int d = 0; d < 3; d++) {
            const int64_t i = nd * 3 + d;
            x[i] = fma(a, p[i], x[i]);
            rv[d] = fma(-a, q[i], r[i]);
            r[i] = rv[d];
        }
        apply_prec(P, nd, rv, zv);
        for (int d = 0; d < 3; d++) {
            z[nd * 3 + d] = zv[d];
            rzn = fma(rv[d], zv[d], rzn);
            rr = fma(rv[d], rv[d], rr);
        }
    }
    rzn = block_sum<WGV>(rzn);
    rr = block_sum<WGV>(rr);
    if (threadIdx.x == 0) {
        part[(pi ? P_RZ1 : P_RZ0) + blockIdx.x] = rzn;
        part[(pi ? P_RR1 : P_RR0) + blockIdx.x] = rr;
    }
}

__global__ void __launch_bounds__(WGV) pcg_dir_kernel(int64_t n, double *__restrict__ p, float *__restrict__ p32,
                                                      const double *__restrict__ z, const double *__restrict__ part,
                                                      int nv, int pi, int iter, int maxiter,
                                                      int32_t *__restrict__ ctl, double *__restrict__ scal)
{
    if (ctl[C_DONE]) return;
    const double rzn = sum_partials<WGV>(part + (pi ? P_RZ1 : P_RZ0), nv);
    const double rzo = sum_partials<WGV>(part + (pi ? P_RZ0 : P_RZ1), nv);
    const double rr = sum_partials<WGV>(part + (pi ? P_RR1 : P_RR0), nv);
    const double beta = rzn / rzo;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl[C_ITERS] = iter;
        scal[S_RR] = rr;
        scal[S_RZ] = rzn;
        if (rr <= scal[S_THR] || iter >= maxiter || !std::isfinite(rr)) ctl[C_STOP] = 1;
    }
    for (int64_t i = (int64_t)blockIdx.x * WGV + threadIdx.x; i < n; i += (int64_t)gridDim.x * WGV) {
        const double pn = fma(beta, p[i], z[i]);
        p[i] = pn;
        if (p32) p32[i] = (float)pn;
    }
}

// x = 0, r = b, z = M r, p = z; partials rz -> P_RZ1, rr -> P_RR1
__global__ void __launch_bounds__(WGV) pcg_init_kernel(int64_t nnodes, const double *__restrict__ b,
                                                       double *__restrict__ x, double *__restrict__ r,
                                                       double *__restrict__ z, double *__restrict__ p,
                                                       float *__restrict__ p32, const Prec P,
                                                       double *__restrict__ part)
{
    double rz = 0.0, rr = 0.0;
    for (int64_t nd = (int64_t)blockIdx.x * WGV + threadIdx.x; nd < nnodes; nd += (int64_t)gridDim.x * WGV) {
        double rv[3], zv[3];
        for (int d = 0; d < 3; d++) rv[d] = b[nd * 3 + d];
        apply_prec(P, nd, rv, zv);
        for (int d = 0; d < 3; d++) {
            const int64_t i = nd * 3 + d;
            x[i] = 0.0;
            r[i] = rv[d];
            z[i] = zv[d];
            p[i] = zv[d];
            if (p32) p32[i] = (float)zv[d];
            rz = fma(rv[d], zv[d], rz);
            rr = fma(rv[d], rv[d], rr);
        }
    }
    rz = block_sum<WGV>(rz);
    rr = block_sum<WGV>(rr);
    if (threadIdx.x == 0) {
        part[P_RZ1 + blockIdx.x] = rz;
        part[P_RR1 + blockIdx.x] = rr;
    }
}

// r = rt; z = M r; with restart also p = z (the search direction starts
// afresh); rz partials into the slot the next iteration reads as "previous".
__global__ void __launch_bounds__(WGV) pcg_rez_kernel(int64_t nnodes, const double *__restrict__ rt,
                                                      double *__restrict__ r, double *__restrict__ z,
                                                      double *__restrict__ p, float *__restrict__ p32, int restart,
                                                      const Prec P, double *__restrict__ part, int slot_pi)
{
    double rz = 0.0;
    for (int64_t nd = (int64_t)blockIdx.x * WGV + threadIdx.x; nd < nnodes; nd += (int64_t)gridDim.x * WGV) {
        double rv[3], zv[3];
        for (int d = 0; d < 3; d++) rv[d] = rt[nd * 3 + d];
        apply_prec(P, nd, rv, zv);
        for (int d = 0; d < 3; d++) {
            const int64_t i = nd * 3 + d;
            r[i] = rv[d];
            z[i] = zv[d];
            if (restart) {
                p[i] = zv[d];
                if (p32) p32[i] = (float)zv[d];
            }
            rz = fma(rv[d], zv[d], rz);
        }
    }
    rz = block_sum<WGV>(rz);
    if (threadIdx.x == 0) part[(slot_pi ? P_RZ1 : P_RZ0) + blockIdx.x] = rz;
}

// One workgroup publishes scalars from partials.
//   mode 0 (after init):        rz, rr = bb from slot 1; recursive threshold; reset control
//   mode 1 (after replacement): rr, bb, kx from the residual partials, rz from the slot
__global__ void __launch_bounds__(WGV) pcg_finalize_kernel(int mode, int nv, int nr, int slot_pi, double thr_fac,
                                                           const double *__restrict__ part,
                                                           int32_t *__restrict__ ctl, double *__restrict__ scal)
{
    if (mode == 0) {
        const double rz = sum_partials<WGV>(part + P_RZ1, nv);
        const double rr = sum_partials<WGV>(part + P_RR1, nv);
        if (threadIdx.x == 0) {
            scal[S_RZ] = rz;
            scal[S_RR] = rr;
            scal[S_BB] = rr;
            scal[S_KX] = 0.0;
            scal[S_THR] = thr_fac * rr;
            ctl[C_ITERS] = 0;
            ctl[C_BREAK] = 0;
            ctl[C_STOP] = 0;
            ctl[C_DONE] = rr == 0.0 ? 1 : 0;
        }
    } else if (mode == 1) {
        const double rr = sum_partials<WGV>(part + P_A, nr);
        const double bb = sum_partials<WGV>(part + P_B, nr);
        const double kx = sum_partials<WGV>(part + P_C, nr);
        const double gap = sum_partials<WGV>(part + P_D, nr);
        if (threadIdx.x == 0) {
            scal[S_RR] = rr;
            scal[S_BB] = bb;
            scal[S_KX] = kx;
            scal[S_GAP] = gap;
        }
    } else {
        const double rz = sum_partials<WGV>(part + (slot_pi ? P_RZ1 : P_RZ0), nv);
        if (threadIdx.x == 0) {
            scal[S_RZ] = rz;
            ctl[C_DONE] = 0;
            ctl[C_STOP] = 0;
        }
    }
}

__global__ void clear_break_kernel(int32_t *__restrict__ ctl) { ctl[C_BREAK] = 0; }

__global__ void fill_f32_kernel(int64_t n, float *__restrict__ x, float v)
{
    for (int64_t i = (int64_t)blockIdx.x * WGV + threadIdx.x; i < n; i += (int64_t)gridDim.x * WGV) x[i] = v;
}

__global__ void fill_kernel(int64_t n, double *__restrict__ x, double v)
{
    for (int64_t i = (int64_t)blockIdx.x * WGV + threadIdx.x; i < n; i += (int64_t)gridDim.x * WGV) x[i] = v;
}

struct Launch {
    int R;                     // block rows per SpMV workgroup
    unsigned grid_rows, grid_vec;
};

Launch pick_launch(const amr_model *m)
{
    Launch L;
    int R = 128;
    while ((m->nnodes + R - 1) / R > MAX_GR) R *= 2;
    L.R = R;
    L.grid_rows = (unsigned)std::max<int64_t>(1, (m->nnodes + R - 1) / R);
    const int64_t nv = (m->nnodes * 3 + WGV - 1) / WGV;
    L.grid_vec = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nv, MAX_GV));
    return L;
}

size_t tile_lds(int R, int ndbl) { return (size_t)(((R + 1) * 4 + 15) & ~15) + (size_t)ndbl * R * 8; }

template <typename VT, typename XT, typename AT = VT>
void launch_pcg_spmv_lanes(amr_model *m, const Launch &L, int lanes, const VT *vals, const XT *px, const double *p,
                           double *q)
{
    int32_t *ctl = (int32_t *)(m->scal + S_NSCAL);
    const size_t lds = tile_lds(L.R, 1);
#define AMR_PCG_SPMV(LN)                                                                                           \
    hipLaunchKernelGGL((pcg_spmv_kernel<LN, VT, XT, AT>), dim3(L.grid_rows), dim3(WGS), lds, m->stream, m->nnodes, \
                       L.R, m->rowptr, m->col, vals, m->pat.nblk, px, p, q, m->mask, m->part, ctl)
    switch (lanes) {
    case 8: AMR_PCG_SPMV(8); break;
    case 16: AMR_PCG_SPMV(16); break;
    case 32: AMR_PCG_SPMV(32); break;
    default: AMR_PCG_SPMV(64); break;
    }
#undef AMR_PCG_SPMV
}

// Autotuner: time each lanes-per-row choice on this model's matrix and keep
// the fastest.  Every choice gives bit-identical results (see row_sum), so
// the choice only affects speed.
template <typename VT, typename XT, typename AT = VT>
int tune_lanes(amr_model *m, const Launch &L, const VT *vals, const XT *px, const double *p, double *q, int *out)
{
    static const int cand[4] = {8, 16, 32, 64};
    hipEvent_t e0, e1;
    AMR_HIP(hipEventCreate(&e0));
    AMR_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int bestL = 64;
    for (int c : cand) {
        launch_pcg_spmv_lanes<VT, XT, AT>(m, L, c, vals, px, p, q);  // warm
        AMR_HIP(hipEventRecord(e0, m->stream));
        for (int rep = 0; rep < 3; rep++) launch_pcg_spmv_lanes<VT, XT, AT>(m, L, c, vals, px, p, q);
        AMR_HIP(hipEventRecord(e1, m->stream));
        AMR_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        AMR_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            bestL = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *out = bestL;
    return AMR_OK;
}

template <typename VT, typename XT, typename AT = VT>
void launch_pcg_spmv(amr_model *m, const Launch &L, const VT *vals, const XT *px, const double *p, double *q)
{
    const int lanes = sizeof(VT) == 8 ? m->lanes64 : (sizeof(AT) == 8 ? m->lanes32d : m->lanes32);
    launch_pcg_spmv_lanes<VT, XT, AT>(m, L, lanes, vals, px, p, q);
}

void launch_residual(amr_model *m, const Launch &L, const double *x, const double *b, double *r, const double *rold)
{
    const size_t lds = tile_lds(L.R, 4);
#define AMR_RES(LN)                                                                                                \
    hipLaunchKernelGGL((residual_kernel<LN>), dim3(L.grid_rows), dim3(WGS), lds, m->stream, m->nnodes, L.R,       \
                       m->rowptr, m->col, m->Kv, m->pat.nblk, x, b, r, rold, m->mask, m->part)
    switch (m->lanes64) {
    case 8: AMR_RES(8); break;
    case 16: AMR_RES(16); break;
    case 32: AMR_RES(32); break;
    default: AMR_RES(64); break;
    }
#undef AMR_RES
}

int read_scalars(amr_model *m, double *hs, int32_t *hc)
{
    AMR_HIP(hipMemcpyAsync(hs, m->scal, sizeof(double) * S_NSCAL, hipMemcpyDeviceToHost, m->stream));
    AMR_HIP(hipMemcpyAsync(hc, m->scal + S_NSCAL, sizeof(int32_t) * C_NCTL, hipMemcpyDeviceToHost, m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    return AMR_OK;
}

}  // namespace

// y = K x (mask_mode 0) or mask .* (K x) (mask_mode 1), FP64, node order.
int amr_spmv_full(amr_model *m, const double *x, double *y, int mask_mode)
{
    if (int rc = amr_autotune(m)) return rc;
    const Launch L = pick_launch(m);
    const size_t lds = tile_lds(L.R, 0);
#define AMR_SPMV(LN)                                                                                               \
    hipLaunchKernelGGL((spmv_kernel<LN>), dim3(L.grid_rows), dim3(WGS), lds, m->stream, m->nnodes, L.R, m->rowptr, \
                       m->col, m->Kv, m->pat.nblk, x, y, m->mask, mask_mode)
    switch (m->lanes64) {
    case 8: AMR_SPMV(8); break;
    case 16: AMR_SPMV(16); break;
    case 32: AMR_SPMV(32); break;
    default: AMR_SPMV(64); break;
    }
#undef AMR_SPMV
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

// Pick the SpMV lanes-per-row for both precisions (once per model).
int amr_autotune(amr_model *m)
{
    if (m->lanes64 && m->lanes32 && m->lanes32d) return AMR_OK;
    const Launch L = pick_launch(m);
    int rc_ = AMR_OK;
    int32_t *ctl = (int32_t *)(m->scal + S_NSCAL);
    AMR_HIP(hipMemsetAsync(ctl, 0, sizeof(int32_t) * C_NCTL, m->stream));
    const int64_t n = m->nnodes * 3;
    hipLaunchKernelGGL(fill_kernel, dim3(L.grid_vec), dim3(WGV), 0, m->stream, n, m->v[3], 1.0);
    hipLaunchKernelGGL(fill_f32_kernel, dim3(L.grid_vec), dim3(WGV), 0, m->stream, n, m->f[0], 1.0f);
    AMR_HIP(hipGetLastError());
    int l64 = 64, l32 = 64, l32d = 64;
    if (!m->lanes64 && (rc_ = tune_lanes<double, double>(m, L, m->Kv, m->v[3], m->v[3], m->v[4], &l64))) return rc_;
    if (!m->lanes32 && (rc_ = tune_lanes<float, float>(m, L, m->Kv32, m->f[0], m->v[3], m->v[4], &l32))) return rc_;
    if (!m->lanes32d &&
        (rc_ = tune_lanes<float, double, double>(m, L, m->Kv32, m->v[3], m->v[3], m->v[4], &l32d)))
        return rc_;
    if (!m->lanes64) m->lanes64 = l64;
    if (!m->lanes32) m->lanes32 = l32;
    if (!m->lanes32d) m->lanes32d = l32d;
    return AMR_OK;
}

// Solve K11 x = b (node order, b zero at prescribed dofs), x FP64.
//
// Residual replacement: when the recursive residual has dropped by 10x
// since the last recomputation (mixed path) or met the tolerance (both
// paths), r = b - K x is recomputed in FP64.  The drift between the two
// residuals measures how well the FP32 operator represents K for this
// model: if it exceeds 0.3 ||r||, the rest of the solve runs with the FP64
// operator (restarting the search direction), so an ill-conditioned model
// still converges to the FP64 solution.
int amr_pcg_solve(amr_model *m, const double *b, double *x, const amr_solve_opts *o, amr_solve_info *info)
{
    const int64_t n = m->nnodes * 3;
    const Launch L = pick_launch(m);
    // mixed: FP32 values; f32x: FP32 search direction and FP32 accumulation
    // (AMR_PREC_MIXED), otherwise FP64 gather and accumulation of the FP32
    // values (AMR_PREC_MIXED_F64ACC)
    bool mixed = o->precision == AMR_PREC_MIXED || o->precision == AMR_PREC_MIXED_F64ACC;
    const bool f32x = o->precision == AMR_PREC_MIXED;
    const int check = std::max(1, std::min(o->check_every > 0 ? o->check_every : 32, 256));
    const double kEps = 2.220446049250313e-16, kFloorC = 64.0;
    // residual drop that triggers a recomputation (mixed path), and the
    // drift/||r|| above which the FP32 operator is abandoned
    const double replace_ratio = o->replace_ratio > 0.0 ? o->replace_ratio : 0.1;
    const double drift_limit = o->drift_limit > 0.0 ? o->drift_limit : 0.3;
    if (int rc = amr_autotune(m)) return rc;
    int32_t *ctl = (int32_t *)(m->scal + S_NSCAL);
    double *r = m->v[1], *z = m->v[2], *p = m->v[3], *q = m->v[4], *rt = m->v[5];
    float *p32 = m->f[0];
    double hs[S_NSCAL];
    int32_t hc[C_NCTL];

    if (m->timing && (int)m->ev.size() < 2 * check) {
        for (hipEvent_t e : m->ev) (void)hipEventDestroy(e);
        m->ev.assign(2 * check, nullptr);
        for (hipEvent_t &e : m->ev) AMR_HIP(hipEventCreate(&e));
    }
    const double tr = 0.5 * o->rtol;
    Prec P{m->dinv, m->dinv32, m->binv, m->binv32, mixed ? 1 : 0, o->precond == AMR_PRECOND_BLOCK_JACOBI ? 1 : 0};
    hipLaunchKernelGGL(pcg_init_kernel, dim3(L.grid_vec), dim3(WGV), 0, m->stream, m->nnodes, b, x, r, z, p,
                       mixed && f32x ? p32 : nullptr, P, m->part);
    hipLaunchKernelGGL(pcg_finalize_kernel, dim3(1), dim3(WGV), 0, m->stream, 0, (int)L.grid_vec, 0, 1, tr * tr,
                       m->part, ctl, m->scal);
    AMR_HIP(hipGetLastError());
    if (int rc = read_scalars(m, hs, hc)) return rc;
    const double bb = hs[S_BB];
    double rr_ref = hs[S_RR], rr_true_prev = INFINITY, relres = 0.0;
    int converged = bb == 0.0, broke = 0, replacements = 0, iters = 0, fell_back = 0, breakdowns = 0;
    int k = 0;
    while (!converged) {
        for (int j = 0; j < check; j++, k++) {
            const int pi = k & 1;
            if (m->timing) AMR_HIP(hipEventRecord(m->ev[2 * j], m->stream));
            if (mixed && f32x) launch_pcg_spmv<float, float>(m, L, m->Kv32, p32, p, q);
            else if (mixed) launch_pcg_spmv<float, double, double>(m, L, m->Kv32, p, p, q);
            else launch_pcg_spmv<double, double>(m, L, m->Kv, p, p, q);
            if (m->timing) AMR_HIP(hipEventRecord(m->ev[2 * j + 1], m->stream));
            hipLaunchKernelGGL(pcg_update_kernel, dim3(L.grid_vec), dim3(WGV), 0, m->stream, m->nnodes, x, r, z, p, q,
                               P, m->part, (int)L.grid_rows, (int)L.grid_vec, pi, ctl);
            hipLaunchKernelGGL(pcg_dir_kernel, dim3(L.grid_vec), dim3(WGV), 0, m->stream, n, p,
                               mixed && f32x ? p32 : nullptr, z, m->part, (int)L.grid_vec, pi, k + 1, o->maxiter, ctl,
                               m->scal);
        }
        AMR_HIP(hipGetLastError());
        if (int rc = read_scalars(m, hs, hc)) return rc;
        if (m->timing) {
            // the SpMVs of the iterations run in this batch executed in full;
            // later launches returned at once on the done flag
            const int real = std::min(hc[C_ITERS] - iters, check);
            for (int j = 0; j < real; j++) {
                float ms = 0.f;
                AMR_HIP(hipEventElapsedTime(&ms, m->ev[2 * j], m->ev[2 * j + 1]));
                m->t_spmv_ms += ms;
                m->n_spmv++;
            }
        }
        iters = hc[C_ITERS];
        broke = hc[C_BREAK];
        k = iters;
        const bool done = hc[C_DONE] != 0 || hc[C_STOP] != 0;
        if (!(done || (mixed && hs[S_RR] <= replace_ratio * replace_ratio * rr_ref))) {
            if (broke || iters >= o->maxiter) break;
            continue;
        }
        // residual replacement: rt = b - K x in FP64, drift ||rt - r||
        const double rr_rec = hs[S_RR];
        launch_residual(m, L, x, b, rt, r);
        hipLaunchKernelGGL(pcg_finalize_kernel, dim3(1), dim3(WGV), 0, m->stream, 1, (int)L.grid_vec,
                           (int)L.grid_rows, 0, 0.0, m->part, ctl, m->scal);
        AMR_HIP(hipGetLastError());
        if (int rc = read_scalars(m, hs, hc)) return rc;
        replacements++;
        const double rr = hs[S_RR];
        relres = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
        const double floor2 = kFloorC * kFloorC * kEps * kEps * hs[S_KX];
        if (rr <= o->rtol * o->rtol * bb || rr <= floor2) {
            converged = 1;
            break;
        }
        if (!std::isfinite(rr) || iters >= o->maxiter) break;
        if (done && !broke && rr > 0.25 * rr_true_prev) break;  // stagnation above both bounds
        if (broke && ++breakdowns > 3) break;
        bool restart = broke != 0;
        if (mixed && hs[S_GAP] > drift_limit * drift_limit * std::max(rr, rr_rec)) {
            mixed = false;  // the FP32 operator is too coarse for this model
            fell_back = 1;
            restart = true;
        }
        const int slot = 1 - (iters & 1);
        P.f32 = mixed ? 1 : 0;
        hipLaunchKernelGGL(pcg_rez_kernel, dim3(L.grid_vec), dim3(WGV), 0, m->stream, m->nnodes, rt, r, z, p,
                           mixed && f32x ? p32 : nullptr, restart ? 1 : 0, P, m->part, slot);
        hipLaunchKernelGGL(pcg_finalize_kernel, dim3(1), dim3(WGV), 0, m->stream, 2, (int)L.grid_vec,
                           (int)L.grid_rows, slot, 0.0, m->part, ctl, m->scal);
        AMR_HIP(hipGetLastError());
        if (broke) {
            hipLaunchKernelGGL(clear_break_kernel, dim3(1), dim3(1), 0, m->stream, ctl);
            broke = 0;
        }
        rr_true_prev = rr;
        rr_ref = rr;
    }
    if (info) {
        info->iterations = iters;
        info->refinements = replacements;
        info->converged = converged;
        info->fallback = fell_back;
        info->relres = relres;
    }
    if (!converged) {
        char buf[256];
        snprintf(buf, sizeof buf, "PCG did not reach rtol=%.3e (relres %.3e after %d iterations, %d replacements%s)",
                 o->rtol, relres, iters, replacements, broke ? ", breakdown" : "");
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
    for (int64_t i = (int64_t)blockIdx.x * WGV + threadIdx.x; i < n; i += (int64_t)gridDim.x * WGV) {
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
    for (int64_t i = (int64_t)blockIdx.x * WGV + threadIdx.x; i < n; i += (int64_t)gridDim.x * WGV)
        u2[i] = mask[i] != 0.0 ? 0.0 : u[i];
}

__global__ void masked_diff_kernel(int64_t n, const double *__restrict__ mask, const double *__restrict__ a,
                                   const double *__restrict__ b, double *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * WGV + threadIdx.x; i < n; i += (int64_t)gridDim.x * WGV)
        out[i] = mask[i] != 0.0 ? a[i] - b[i] : 0.0;
}

__global__ void combine_kernel(int64_t n, const double *__restrict__ mask, const double *__restrict__ x,
                               double *__restrict__ u)
{
    for (int64_t i = (int64_t)blockIdx.x * WGV + threadIdx.x; i < n; i += (int64_t)gridDim.x * WGV)
        if (mask[i] != 0.0) u[i] = x[i];
}

__global__ void vadd_kernel(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                            double *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * WGV + threadIdx.x; i < n; i += (int64_t)gridDim.x * WGV) out[i] = a[i] + b[i];
}

// ΔUa += ΔUi; R = mask .* (ΔFex - ΔFin); per-workgroup max|R|, max|ΔUi|,
// max|ΔUa| and a NaN-in-ΔFin flag
__global__ void __launch_bounds__(WGV) nr_post_kernel(int64_t n, const double *__restrict__ mask,
                                                      const double *__restrict__ dFex,
                                                      const double *__restrict__ dFin,
                                                      const double *__restrict__ dUi, double *__restrict__ dUa,
                                                      double *__restrict__ R, double *__restrict__ part)
{
    double mr = 0.0, mi = 0.0, ma = 0.0, bad = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * WGV + threadIdx.x; i < n; i += (int64_t)gridDim.x * WGV) {
        const double ua = dUa[i] + dUi[i];
        dUa[i] = ua;
        const double fi = dFin[i];
        const double r = mask[i] != 0.0 ? dFex[i] - fi : 0.0;
        R[i] = r;
        if (fi != fi) bad = 1.0;
        mr = fmax(mr, fabs(r));
        mi = fmax(mi, fabs(dUi[i]));
        ma = fmax(ma, fabs(ua));
    }
    mr = block_max<WGV>(mr);
    mi = block_max<WGV>(mi);
    ma = block_max<WGV>(ma);
    bad = block_max<WGV>(bad);
    if (threadIdx.x == 0) {
        part[P_A + blockIdx.x] = mr;
        part[P_B + blockIdx.x] = mi;
        part[P_C + blockIdx.x] = ma;
        part[P_D + blockIdx.x] = bad;
    }
}

__global__ void __launch_bounds__(WGV) nr_finalize_kernel(int nv, const double *__restrict__ part,
                                                          int32_t *__restrict__ ctl, double *__restrict__ scal)
{
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nv; i += WGV)
        for (int k = 0; k < 4; k++) v[k] = fmax(v[k], part[P_A + k * 4096 + i]);
    for (int k = 0; k < 4; k++) v[k] = block_max<WGV>(v[k]);
    if (threadIdx.x == 0) {
        scal[S_RES] = v[0];
        scal[S_ERRN] = v[1];
        scal[S_ERRD] = v[2];
        ctl[C_NAN] = v[3] != 0.0;
    }
}

__global__ void commit_kernel(int64_t n, double acr, const double *__restrict__ dUa, const double *__restrict__ dFin,
                              const double *__restrict__ R, double *__restrict__ U, double *__restrict__ F,
                              double *__restrict__ Rc)
{
    for (int64_t i = (int64_t)blockIdx.x * WGV + threadIdx.x; i < n; i += (int64_t)gridDim.x * WGV) {
        U[i] += dUa[i];
        F[i] += dFin[i];
        Rc[i] = (1.0 - acr) * Rc[i] + R[i];
    }
}

unsigned vec_grid(const amr_model *m)
{
    const int64_t nv = (m->nnodes * 3 + WGV - 1) / WGV;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(nv, MAX_GV));
}

}  // namespace

int amr_vec_inc_begin(amr_model *m, double dT, double acr, int add)
{
    double **v = m->sv;
    hipLaunchKernelGGL(inc_begin_kernel, dim3(vec_grid(m)), dim3(WGV), 0, m->stream, m->nnodes * 3, dT, acr, add,
                       v[SV_UEX], v[SV_FEX], v[SV_RC], v[SV_DFEX], v[SV_R], v[SV_DUA], v[SV_DUI]);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

int amr_vec_fill(amr_model *m, double *x, double val)
{
    hipLaunchKernelGGL(fill_kernel, dim3(vec_grid(m)), dim3(WGV), 0, m->stream, m->nnodes * 3, x, val);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

// b = mask .* (R - K (prescribed part of dUi)); u2/ky are scratch
int amr_vec_solve_prep(amr_model *m, const double *dUi, const double *R, double *u2, double *ky, double *b)
{
    const int64_t n = m->nnodes * 3;
    hipLaunchKernelGGL(prescribed_part_kernel, dim3(vec_grid(m)), dim3(WGV), 0, m->stream, n, m->mask, dUi, u2);
    AMR_HIP(hipGetLastError());
    int rc = amr_spmv_full(m, u2, ky, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(masked_diff_kernel, dim3(vec_grid(m)), dim3(WGV), 0, m->stream, n, m->mask, R, ky, b);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

int amr_vec_combine(amr_model *m, const double *x, double *dUi)
{
    hipLaunchKernelGGL(combine_kernel, dim3(vec_grid(m)), dim3(WGV), 0, m->stream, m->nnodes * 3, m->mask, x, dUi);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

int amr_vec_add(amr_model *m, const double *a, const double *b, double *out)
{
    hipLaunchKernelGGL(vadd_kernel, dim3(vec_grid(m)), dim3(WGV), 0, m->stream, m->nnodes * 3, a, b, out);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

int amr_vec_nr_post(amr_model *m, double *res, double *dui_max, double *dua_max, int *nan_flag)
{
    double **v = m->sv;
    int32_t *ctl = (int32_t *)(m->scal + S_NSCAL);
    const unsigned g = vec_grid(m);
    hipLaunchKernelGGL(nr_post_kernel, dim3(g), dim3(WGV), 0, m->stream, m->nnodes * 3, m->mask, v[SV_DFEX],
                       v[SV_DFIN], v[SV_DUI], v[SV_DUA], v[SV_R], m->part);
    hipLaunchKernelGGL(nr_finalize_kernel, dim3(1), dim3(WGV), 0, m->stream, (int)g, m->part, ctl, m->scal);
    AMR_HIP(hipGetLastError());
    double hs[S_NSCAL];
    int32_t hc[C_NCTL];
    if (int rc = read_scalars(m, hs, hc)) return rc;
    *res = hs[S_RES];
    *dui_max = hs[S_ERRN];
    *dua_max = hs[S_ERRD];
    *nan_flag = hc[C_NAN];
    return AMR_OK;
}

int amr_vec_commit(amr_model *m)
{
    double **v = m->sv;
    hipLaunchKernelGGL(commit_kernel, dim3(vec_grid(m)), dim3(WGV), 0, m->stream, m->nnodes * 3, m->acr, v[SV_DUA],
                       v[SV_DFIN], v[SV_R], v[SV_U], v[SV_F], v[SV_RC]);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}
