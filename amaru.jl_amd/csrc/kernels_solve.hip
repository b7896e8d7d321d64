// Solver kernels for gfx950: node-block (3x3) SpMV and Jacobi-PCG, replacing
// the UMFPACK solve of solve_system! (src/solver.jl:42-43), plus the vector
// kernels of the device-resident stage driver.
//
// Vectors live in node order (3 dofs per node).  The operator is K with the
// prescribed dofs masked out: y = mask .* (K x).  With x zero at prescribed
// dofs this is K11 x on the unknowns, so one resident copy of K serves K11,
// K12 U2 (right-hand side) and K21 U1 + K22 U2 (reactions).
//
// Pools.  The PCG kernels work on a pool of independent systems (models) at
// once: a descriptor per model holds its matrix, vectors, partials and
// scalars, and every launch walks a work list of (model, row tile) or
// (model, vector block) entries, so one launch per PCG phase advances every
// model of the pool.  Each model's tile and block decomposition depends on
// that model alone, and each model's scalars are reduced from its own
// partials in a fixed order, so a model's iterates are bit-identical whether
// it is solved alone (a pool of one, which is what amr_pcg_solve does) or
// together with any others.  Models join and leave the pool between steps
// (continuous batching at the granularity of one step of check_every
// iterations).
//
// Reductions need no atomics and no fences: a reducing kernel writes one
// partial per workgroup, and every workgroup of the next kernel that needs
// the scalar sums those partials itself in index order (kernel boundaries
// make them visible).
//
// Precision.  Plain path: FP64 PCG.  Mixed paths: the same CG recurrences
// with FP64 vectors, FP64 dot products and FP64 solution update, but the
// SpMV reads FP32 values (with an FP32 copy of the search direction and FP32
// accumulation for AMR_PREC_MIXED, FP64 gather and accumulation for
// AMR_PREC_MIXED_F64ACC) and the preconditioner is applied in FP32.  The
// FP64 residual r = b - K x is recomputed with the FP64 matrix whenever the
// recursive residual has dropped 10x since the last recomputation and at the
// end (residual replacement), which keeps the iteration converging to the
// FP64 solution without restarting the Krylov process; a model whose FP32
// operator drifts too far from K falls back to the FP64 operator.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "amr_internal.h"

namespace {

constexpr int WGV = 256;  // vector kernels
constexpr int WGS = 512;  // SpMV kernels
constexpr int MAX_GR = 8192, MAX_GV = 1024;

// scalar and control slots of one model: S_NSCAL doubles, then C_NCTL ints
// S_RZP0/1: r.z published per parity slot (the sum of that slot's partials)
enum { S_RZ = 0, S_RR, S_BB, S_THR, S_KX, S_RES, S_ERRN, S_ERRD, S_GAP, S_RZP0, S_RZP1, S_NSCAL = 16 };
__device__ inline int rz_pub(int slot) { return slot ? S_RZP1 : S_RZP0; }
// C_DONE is read by every workgroup at kernel entry and written only where
// all workgroups of the model in the running launch agree on it
// (breakdown); the convergence test writes C_STOP, which the next SpMV
// promotes to C_DONE, so no workgroup ever sees a flag raised during its
// own launch.  C_ITERS is advanced by the SpMV's first tile and read by the
// kernels after it.
enum { C_DONE = 0, C_ITERS, C_BREAK, C_NAN, C_STOP, C_NCTL = 16 };
constexpr int SEG_DBL = 32;  // doubles per model in a pool's scalar array
static_assert(S_NSCAL + C_NCTL / 2 <= SEG_DBL, "scalar block");
// partial buffers inside m->part
enum { P_PQ = 0, P_RZ0 = 8192, P_RZ1 = 9216, P_RR0 = 10240, P_RR1 = 11264, P_CY0 = 12288, P_CY1 = 12800, P_A = 13312,
       P_B = 21504, P_C = 29696, P_D = 37888, P_END = 46080 };
static_assert(P_END == AMR_PART_DBL, "partials layout");
constexpr int MAX_CW = 512;  // coarse-product workgroups per model (4 coarse rows each)

// Pointers taken from a descriptor are generic; device buffers are all in
// the global address space, and saying so lets the compiler use global
// (not flat) memory instructions.
template <typename T>
__device__ inline const __attribute__((address_space(1))) T *gl(const T *p)
{
    return (const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ inline __attribute__((address_space(1))) T *glw(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

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
    // up to 16 loads of a thread issued before its additions; the additions
    // keep the order i, i+NT, ...
    const auto *Pg = gl(P);
    for (int i = threadIdx.x; i < n; i += 16 * NT) {
        double a[16];
#pragma unroll
        for (int k = 0; k < 16; k++) a[k] = i + k * NT < n ? Pg[i + k * NT] : 0.0;
#pragma unroll
        for (int k = 0; k < 16; k++)
            if (i + k * NT < n) v += a[k];
    }
    return block_sum<NT>(v);
}

// One model of a pool, as the kernels see it.
struct SegDesc {
    const int32_t *rowptr, *col;
    const int16_t *col16;         // column - row per block when every offset fits 16 bits, else null
    const double *Kv;
    const float *Kv32;
    const double *mask;
    const double *dinv, *binv;    // Jacobi / nodal block-Jacobi inverses
    const float *dinv32, *binv32;
    const double *b;              // right-hand side (node order, 0 at prescribed dofs)
    double *x, *r, *z, *p, *q, *rt;
    float *p32;                   // FP32 copy of p (AMR_PREC_MIXED)
    double *part, *scal;
    int32_t *ctl;
    int64_t nnodes, nblk;
    double thr_fac;               // recursive-residual stop: rr <= thr_fac * bb
    int R, ntiles, nvblk, maxiter;
    int f32;                      // preconditioner applied in FP32
    int block;                    // nodal block Jacobi
    int wp32;                     // keep p32 up to date
    int tl;                       // two-level: add the coarse correction Phi Kc^-1 Phi^T r
    // two-level coarse space
    int nagg, nc, ncw;
    const int32_t *agg, *aggptr, *aggnodes;
    const double *cent, *X, *Kinv;
    double *cvec, *yvec;
};

// r.z with z = M r: the Jacobi part's partials plus, two-level, the coarse
// part c.y = (Phi^T r).(Kc^-1 Phi^T r)
__device__ inline double sum_rz(const SegDesc &d, int slot)
{
    double v = sum_partials<WGV>(d.part + (slot ? P_RZ1 : P_RZ0), d.nvblk);
    if (d.tl) v += sum_partials<WGV>(d.part + (slot ? P_CY1 : P_CY0), d.ncw);
    return v;
}

// coarse correction at dof c of node nd: (R_nd y_agg(nd))_c, masked
__device__ inline double coarse_term(const SegDesc &d, int64_t nd, int c)
{
    const int32_t a = gl(d.agg)[nd];
    const auto *y = gl(d.yvec) + (int64_t)a * 6;
    const auto *X = gl(d.X), *cent = gl(d.cent);
    const double dx = X[nd * 3] - cent[a * 3], dy = X[nd * 3 + 1] - cent[a * 3 + 1],
                 dz = X[nd * 3 + 2] - cent[a * 3 + 2];
    double v;
    if (c == 0) v = fma(-dy, y[5], fma(dz, y[4], y[0]));
    else if (c == 1) v = fma(dx, y[5], fma(-dz, y[3], y[1]));
    else v = fma(-dx, y[4], fma(dy, y[3], y[2]));
    return gl(d.mask)[nd * 3 + c] != 0.0 ? v : 0.0;
}

// Preconditioner: Jacobi (inverse diagonal) or 3x3 nodal block Jacobi
// (inverse of each node's diagonal block restricted to its free dofs), in
// FP64 or applied in FP32 (mixed paths).
__device__ inline void apply_prec(const SegDesc &P, int64_t node, const double r[3], double z[3])
{
    const int64_t i = node * 3;
    if (P.block) {
        if (P.f32) {
            const auto *B = gl(P.binv32) + node * 9;
            const float r0 = (float)r[0], r1 = (float)r[1], r2 = (float)r[2];
            for (int d = 0; d < 3; d++) z[d] = (double)fmaf(B[d * 3 + 2], r2, fmaf(B[d * 3 + 1], r1, B[d * 3] * r0));
        } else {
            const auto *B = gl(P.binv) + node * 9;
            for (int d = 0; d < 3; d++) z[d] = fma(B[d * 3 + 2], r[2], fma(B[d * 3 + 1], r[1], B[d * 3] * r[0]));
        }
    } else if (P.f32) {
        for (int d = 0; d < 3; d++) z[d] = (double)(gl(P.dinv32)[i + d] * (float)r[d]);
    } else {
        for (int d = 0; d < 3; d++) z[d] = gl(P.dinv)[i + d] * r[d];
    }
}

// ------------------------------------------------------------------ SpMV
//
// Row tiles: a workgroup owns block rows [t*R, (t+1)*R) of one model (R a
// power of two fixed per model), stages their row pointers in LDS and walks
// them with L lanes per row.  Each row's sum is formed over 64 virtual slots
// (slot j takes blocks rb+j, rb+j+64, ... in order) reduced by one fixed
// binary tree, whatever L is: a lane holds 64/L slots and performs the tree
// levels above L itself, the levels below L are cross-lane shuffles.
// Per-row dot contributions go to LDS and are reduced by a fixed tree too.
// So neither the row sums nor the partials depend on L, and L can be
// chosen per model for speed alone.


// The nine values of block k, converted to the accumulation type.
typedef double f64x2 __attribute__((ext_vector_type(2)));

template <typename VT, typename AT> struct BlockLoad;

// FP64 values (kv_at layout): four 2-wide loads and one scalar load
template <typename AT> struct BlockLoad<double, AT> {
    __device__ static inline void load(const double *__restrict__ vals, int64_t k, AT v[9])
    {
        const auto *base = gl(vals) + (k >> 6) * 576;
        const auto *pr = gl(reinterpret_cast<const f64x2 *>(vals + (k >> 6) * 576)) + (k & 63);
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const f64x2 w = pr[p * 64];
            v[2 * p] = (AT)w.x;
            v[2 * p + 1] = (AT)w.y;
        }
        v[8] = (AT)base[512 + (k & 63)];
    }
};

// FP32 values (kv32_at layout): two 4-wide loads and one scalar load
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <typename AT> struct BlockLoad<float, AT> {
    __device__ static inline void load(const float *__restrict__ vals, int64_t k, AT v[9])
    {
        const auto *base = gl(vals) + (k >> 6) * 576;
        const auto *pq = gl(reinterpret_cast<const f32x4 *>(vals + (k >> 6) * 576)) + (k & 63);
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const f32x4 w = pq[q * 64];
            v[4 * q] = (AT)w.x;
            v[4 * q + 1] = (AT)w.y;
            v[4 * q + 2] = (AT)w.z;
            v[4 * q + 3] = (AT)w.w;
        }
        v[8] = (AT)base[512 + (k & 63)];
    }
};

template <typename VT, typename AT>
__device__ inline void load_block(const VT *__restrict__ vals, int64_t k, AT v[9])
{
    BlockLoad<VT, AT>::load(vals, k, v);
}

// Column of block k: col[k], or with 16-bit column offsets row + col16[k].
template <typename CT>
__device__ inline int64_t block_col(const CT *__restrict__ col, int64_t k, int64_t row);
template <>
__device__ inline int64_t block_col<int32_t>(const int32_t *__restrict__ col, int64_t k, int64_t)
{
    return gl(col)[k];
}
template <>
__device__ inline int64_t block_col<int16_t>(const int16_t *__restrict__ col, int64_t k, int64_t row)
{
    return row + gl(col)[k];
}

template <int L, bool ABS, typename VT, typename XT, typename AT, typename CT = int32_t>
__device__ inline void row_sum(int32_t rb, int32_t re, int sub, const CT *__restrict__ col,
                               const VT *__restrict__ vals, const XT *__restrict__ x, AT y[3], int64_t row = 0)
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
                const int64_t c = block_col<CT>(col, k, row);
                const auto *xc = gl(x) + c * 3;
                const AT x0 = (AT)xc[0], x1 = (AT)xc[1], x2 = (AT)xc[2];
                AT v[9];
                load_block<VT, AT>(vals, k, v);
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

// row_sum<L, false> (y = K x) and row_sum<L, true> (a = |K| |x|) of the
// same row in one pass over its blocks: each sum has the same slots and the
// same tree as alone, so the same bits.
template <int L, typename VT, typename XT, typename AT, typename CT = int32_t>
__device__ inline void row_sum_pair(int32_t rb, int32_t re, int sub, const CT *__restrict__ col,
                                    const VT *__restrict__ vals, const XT *__restrict__ x, AT y[3], AT a[3],
                                    int64_t row = 0)
{
    constexpr int NS = 64 / L;
    AT acc[NS][3], abs_[NS][3];
#pragma unroll
    for (int t = 0; t < NS; t++)
#pragma unroll
        for (int d = 0; d < 3; d++) acc[t][d] = abs_[t][d] = (AT)0;
    for (int32_t base = rb; base < re; base += 64) {
#pragma unroll
        for (int t = 0; t < NS; t++) {
            const int32_t k = base + sub + L * t;
            if (k < re) {
                const int64_t c = block_col<CT>(col, k, row);
                const auto *xc = gl(x) + c * 3;
                const AT x0 = (AT)xc[0], x1 = (AT)xc[1], x2 = (AT)xc[2];
                AT v[9];
                load_block<VT, AT>(vals, k, v);
                acc[t][0] = fma(v[2], x2, fma(v[1], x1, fma(v[0], x0, acc[t][0])));
                acc[t][1] = fma(v[5], x2, fma(v[4], x1, fma(v[3], x0, acc[t][1])));
                acc[t][2] = fma(v[8], x2, fma(v[7], x1, fma(v[6], x0, acc[t][2])));
                abs_[t][0] += fabs(v[0] * x0) + fabs(v[1] * x1) + fabs(v[2] * x2);
                abs_[t][1] += fabs(v[3] * x0) + fabs(v[4] * x1) + fabs(v[5] * x2);
                abs_[t][2] += fabs(v[6] * x0) + fabs(v[7] * x1) + fabs(v[8] * x2);
            }
        }
    }
#pragma unroll
    for (int off = 32; off >= L; off >>= 1) {
#pragma unroll
        for (int t = 0; t < off / L; t++)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                acc[t][d] = acc[t][d] + acc[t + off / L][d];
                abs_[t][d] = abs_[t][d] + abs_[t + off / L][d];
            }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        AT u = acc[0][d], w = abs_[0][d];
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) {
            u += __shfl_xor(u, off, L);
            w += __shfl_xor(w, off, L);
        }
        y[d] = u;
        a[d] = w;
    }
}

// Smallest power of two >= R: LDS reduction arrays are that long, zero
// padded, so the tree is the same fixed binary tree for any R.
__host__ __device__ inline int pow2ceil(int R)
{
    int p = 1;
    while (p < R) p <<= 1;
    return p;
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

// lds_tree_sum's tree (R a power of two, 64 <= R <= 1024) done by the first
// wave alone, after the barrier that completes s: lane i reduces
// s[i + 64 m] over m in registers (the levels h >= 64, same pairs in the
// same order), then the levels h = 32..1 by shuffles.  The result is in lane
// 0; the other waves need not wait.
__device__ inline double wave_tree_sum(const double *s, int R)
{
    const int M = R / 64;
    double v[16];
#pragma unroll
    for (int m = 0; m < 16; m++) v[m] = m < M ? s[threadIdx.x + 64 * m] : 0.0;
#pragma unroll
    for (int H = 8; H > 0; H >>= 1)
        if (H < M)
#pragma unroll
            for (int m = 0; m < H; m++) v[m] = v[m] + v[m + H];
    double a = v[0];
#pragma unroll
    for (int h = 32; h > 0; h >>= 1) a = a + __shfl_down(a, h, 64);
    return a;
}

// Dynamic LDS layout of the row-tile kernels: R+1 row pointers, then up to
// four double arrays of pow2ceil(R) entries (sized for the largest R of the
// launch).
__device__ inline int32_t *tile_rowptr(char *sm) { return (int32_t *)sm; }
__device__ inline double *tile_dbl(char *sm, int R, int j)
{
    const int off = ((R + 1) * 4 + 15) & ~15;
    return (double *)(sm + off) + (int64_t)j * pow2ceil(R);
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8).  Entry t(b) gives each XCD one contiguous run of the work
// list, so the slice of x a run of row tiles gathers stays in that XCD's L2.
__device__ inline int64_t xcd_tile(int64_t b, int64_t G)
{
    const int64_t x = b & 7, i = b >> 3, q = G >> 3, rm = G & 7;
    return x * q + (x < rm ? x : rm) + i;
}

__device__ inline int load_rows(const int32_t *__restrict__ rowptr, int64_t nrows, int R, int64_t tile,
                                int32_t *srp, int64_t &r0)
{
    r0 = tile * R;
    const int nr = (int)std::min<int64_t>(R, nrows - r0);
    for (int i = threadIdx.x; i <= nr; i += blockDim.x) srp[i] = gl(rowptr)[r0 + i];
    __syncthreads();
    return nr;
}

// y = K x (mode 0) or mask .* (K x) (mode 1), FP64, one model
template <int L>
__global__ void __launch_bounds__(WGS) spmv_kernel(int64_t nrows, int R, const int32_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ col, const double *__restrict__ vals,
                                                   const double *__restrict__ x, double *__restrict__ y,
                                                   const double *__restrict__ mask, int mode)
{
    extern __shared__ __align__(16) char sm[];
    int32_t *srp = tile_rowptr(sm);
    int64_t r0;
    const int nr = load_rows(rowptr, nrows, R, xcd_tile(blockIdx.x, gridDim.x), srp, r0);
    constexpr int RPW = WGS / L;
    const int sub = threadIdx.x % L;
    for (int i = threadIdx.x / L; i < nr; i += RPW) {
        double yy[3];
        row_sum<L, false, double, double, double>(srp[i], srp[i + 1], sub, col, vals, x, yy);
        if (sub == 0) {
            const int64_t row = r0 + i;
            for (int d = 0; d < 3; d++)
                y[row * 3 + d] = (mode == 1 && mask[row * 3 + d] == 0.0) ? 0.0 : yy[d];
        }
    }
}

// rt = mask .* (b - K x) in FP64; per-tile partials ||rt||^2, ||b||^2,
// || |K||x| + |b| ||^2 and the drift ||rt - r||^2 (r: recursive residual)
template <int L>
__global__ void __launch_bounds__(WGS) residual_kernel(const SegDesc *__restrict__ D, const int2 *__restrict__ tl)
{
    const int2 te = tl[xcd_tile(blockIdx.x, gridDim.x)];
    const SegDesc &d = D[te.x];
    const int R = d.R;
    extern __shared__ __align__(16) char sm[];
    int32_t *srp = tile_rowptr(sm);
    double *s_rr = tile_dbl(sm, R, 0), *s_bb = tile_dbl(sm, R, 1), *s_kx = tile_dbl(sm, R, 2),
           *s_gap = tile_dbl(sm, R, 3);
    const int RP = pow2ceil(R);
    for (int i = threadIdx.x; i < RP; i += WGS) s_rr[i] = s_bb[i] = s_kx[i] = s_gap[i] = 0.0;
    int64_t r0;
    const int nr = load_rows(d.rowptr, d.nnodes, R, te.y, srp, r0);
    constexpr int RPW = WGS / L;
    const int sub = threadIdx.x % L;
    const auto *__restrict__ b = gl(d.b);
    const auto *__restrict__ mask = gl(d.mask);
    for (int i = threadIdx.x / L; i < nr; i += RPW) {
        double y[3], a[3];
        if (d.col16)
            row_sum_pair<L, double, double, double, int16_t>(srp[i], srp[i + 1], sub, d.col16, d.Kv, d.x, y, a,
                                                             r0 + i);
        else
            row_sum_pair<L, double, double, double>(srp[i], srp[i + 1], sub, d.col, d.Kv, d.x, y, a);
        if (sub == 0) {
            double rr = 0.0, bb = 0.0, kx = 0.0, gap = 0.0;
            for (int c = 0; c < 3; c++) {
                const int64_t k = (r0 + i) * 3 + c;
                const bool fr = mask[k] != 0.0;
                const double ri = fr ? b[k] - y[c] : 0.0;
                const double bi = fr ? b[k] : 0.0;
                const double ai = fr ? a[c] + fabs(b[k]) : 0.0;
                glw(d.rt)[k] = ri;
                rr = fma(ri, ri, rr);
                bb = fma(bi, bi, bb);
                kx = fma(ai, ai, kx);
                const double g = ri - gl(d.r)[k];
                gap = fma(g, g, gap);
            }
            s_rr[i] = rr;
            s_bb[i] = bb;
            s_kx[i] = kx;
            s_gap[i] = gap;
        }
    }
    __syncthreads();
    const double trr = lds_tree_sum(s_rr, RP), tbb = lds_tree_sum(s_bb, RP), tkx = lds_tree_sum(s_kx, RP),
                 tgap = lds_tree_sum(s_gap, RP);
    if (threadIdx.x == 0) {
        glw(d.part)[P_A + te.y] = trr;
        glw(d.part)[P_B + te.y] = tbb;
        glw(d.part)[P_C + te.y] = tkx;
        glw(d.part)[P_D + te.y] = tgap;
    }
}

// ------------------------------------------------------------------- PCG
//
// Iteration k of a model (parity pi = k % 2), three kernels:
//   pcg_spmv:   q = A p, partial p.q                          -> P_PQ
//   pcg_update: a = rz_k / pq; x += a p; r -= a q; z = M r;
//               partials rz_{k+1}, rr                         -> P_RZ[pi], P_RR[pi]
//   pcg_dir:    b = rz_{k+1} / rz_k; p = z + b p; convergence test;
//               publishes rz_{k+1} in scalar slot rz_pub(pi)
// rz_k is the sum of P_RZ[1-pi], published in rz_pub(1-pi) by iteration
// k-1's pcg_dir, by the initialisation or by a residual replacement.

template <typename VT>
__device__ inline const VT *seg_vals(const SegDesc &d);
template <>
__device__ inline const double *seg_vals<double>(const SegDesc &d) { return d.Kv; }
template <>
__device__ inline const float *seg_vals<float>(const SegDesc &d) { return d.Kv32; }
template <typename XT>
__device__ inline const XT *seg_px(const SegDesc &d);
template <>
__device__ inline const double *seg_px<double>(const SegDesc &d) { return d.p; }
template <>
__device__ inline const float *seg_px<float>(const SegDesc &d) { return d.p32; }

template <int L, typename VT, typename XT, typename AT>
__global__ void __launch_bounds__(WGS) pcg_spmv_kernel(const SegDesc *__restrict__ D, const int2 *__restrict__ tl)
{
    const int2 te = tl[xcd_tile(blockIdx.x, gridDim.x)];
    if (te.x < 0) return;  // padding of a captured step's list
    const SegDesc &d = D[te.x];
    auto *ctl = glw(d.ctl);
    if (ctl[C_DONE]) return;
    if (ctl[C_STOP]) {
        // converged in the previous iteration: every workgroup leaves, and
        // the first tile raises C_DONE for the kernels that follow
        if (te.y == 0 && threadIdx.x == 0) ctl[C_DONE] = 1;
        return;
    }
    if (te.y == 0 && threadIdx.x == 0) ctl[C_ITERS] = ctl[C_ITERS] + 1;
    const int R = d.R;
    extern __shared__ __align__(16) char sm[];
    int32_t *srp = tile_rowptr(sm);
    double *sdot = tile_dbl(sm, R, 0);
    static_assert(sizeof(AT) >= sizeof(VT), "accumulate at least in the storage precision");
    const int RP = pow2ceil(R);
    for (int i = threadIdx.x; i < RP; i += WGS) sdot[i] = 0.0;
    int64_t r0;
    const int nr = load_rows(d.rowptr, d.nnodes, R, te.y, srp, r0);
    constexpr int RPW = WGS / L;
    const int sub = threadIdx.x % L;
    const VT *vals = seg_vals<VT>(d);
    const XT *px = seg_px<XT>(d);
    for (int i = threadIdx.x / L; i < nr; i += RPW) {
        // lanes 0..2 of the row group fetch the row's mask and p entries
        // before its blocks, so the row's tail waits on no load
        const int64_t k3 = (r0 + i) * 3 + (sub < 3 ? sub : 0);
        double mk = 0.0, pk = 0.0;
        if (sub < 3) {
            mk = gl(d.mask)[k3];
            pk = gl(d.p)[k3];
        }
        AT y[3];
        if (d.col16)
            row_sum<L, false, VT, XT, AT, int16_t>(srp[i], srp[i + 1], sub, d.col16, vals, px, y, r0 + i);
        else
            row_sum<L, false, VT, XT, AT>(srp[i], srp[i + 1], sub, d.col, vals, px, y);
        // lane e < 3 stores q of component e; lane 0 forms the row's p.q,
        // fma(p2, q2, fma(p1, q1, fma(p0, q0, 0)))
        const double qk = mk != 0.0 ? (double)(sub == 0 ? y[0] : sub == 1 ? y[1] : y[2]) : 0.0;
        if (sub < 3) glw(d.q)[k3] = qk;
        const double q1 = __shfl_down(qk, 1, L), q2 = __shfl_down(qk, 2, L);
        const double p1 = __shfl_down(pk, 1, L), p2 = __shfl_down(pk, 2, L);
        if (sub == 0) sdot[i] = fma(p2, q2, fma(p1, q1, fma(pk, qk, 0.0)));
    }
    __syncthreads();
    if (RP <= 1024) {
        if (threadIdx.x < 64) {
            const double pq = wave_tree_sum(sdot, RP);
            if (threadIdx.x == 0) glw(d.part)[P_PQ + te.y] = pq;
        }
    } else {
        const double pq = lds_tree_sum(sdot, RP);
        if (threadIdx.x == 0) glw(d.part)[P_PQ + te.y] = pq;
    }
}

__global__ void __launch_bounds__(WGV) pcg_update_kernel(const SegDesc *__restrict__ D, const int2 *__restrict__ vl)
{
    const int2 ve = vl[blockIdx.x];
    if (ve.x < 0) return;
    const SegDesc &d = D[ve.x];
    auto *ctl = glw(d.ctl);
    if (ctl[C_DONE]) return;
    const int pi = (ctl[C_ITERS] - 1) & 1;
    auto *__restrict__ x = glw(d.x);
    auto *__restrict__ r = glw(d.r);
    auto *__restrict__ z = glw(d.z);
    const auto *__restrict__ p = gl(d.p);
    const auto *__restrict__ q = gl(d.q);
    const int64_t stride = (int64_t)d.nvblk * WGV;
    // the first node's operands are loaded before the partial sums, so that
    // their latency overlaps the reductions
    int64_t nd = (int64_t)ve.y * WGV + threadIdx.x;
    double pv[3] = {0.0, 0.0, 0.0}, qv[3] = {0.0, 0.0, 0.0}, xv[3] = {0.0, 0.0, 0.0}, rv0[3] = {0.0, 0.0, 0.0};
    if (nd < d.nnodes)
        for (int c = 0; c < 3; c++) {
            pv[c] = p[nd * 3 + c];
            qv[c] = q[nd * 3 + c];
            xv[c] = x[nd * 3 + c];
            rv0[c] = r[nd * 3 + c];
        }
    const double pq = sum_partials<WGV>(d.part + P_PQ, d.ntiles);
    const double rz = glw(d.scal)[rz_pub(1 - pi)];
    if (!(pq > 0.0) || !std::isfinite(pq)) {
        if (ve.y == 0 && threadIdx.x == 0) {
            ctl[C_BREAK] = 1;
            ctl[C_DONE] = 1;
        }
        return;
    }
    const double a = rz / pq;
    double rzn = 0.0, rr = 0.0;
    // one node per pass; the dot contributions in node order
    while (nd < d.nnodes) {
        double rv[3], zv[3];
        for (int c = 0; c < 3; c++) {
            const int64_t i = nd * 3 + c;
            x[i] = fma(a, pv[c], xv[c]);
            rv[c] = fma(-a, qv[c], rv0[c]);
            r[i] = rv[c];
        }
        apply_prec(d, nd, rv, zv);
        for (int c = 0; c < 3; c++) {
            z[nd * 3 + c] = zv[c];
            rzn = fma(rv[c], zv[c], rzn);
            rr = fma(rv[c], rv[c], rr);
        }
        nd += stride;
        if (nd < d.nnodes)
            for (int c = 0; c < 3; c++) {
                pv[c] = p[nd * 3 + c];
                qv[c] = q[nd * 3 + c];
                xv[c] = x[nd * 3 + c];
                rv0[c] = r[nd * 3 + c];
            }
    }
    rzn = block_sum<WGV>(rzn);
    rr = block_sum<WGV>(rr);
    if (threadIdx.x == 0) {
        glw(d.part)[(pi ? P_RZ1 : P_RZ0) + ve.y] = rzn;
        glw(d.part)[(pi ? P_RR1 : P_RR0) + ve.y] = rr;
    }
}

__global__ void __launch_bounds__(WGV) pcg_dir_kernel(const SegDesc *__restrict__ D, const int2 *__restrict__ vl)
{
    const int2 ve = vl[blockIdx.x];
    if (ve.x < 0) return;
    const SegDesc &d = D[ve.x];
    auto *ctl = glw(d.ctl);
    if (ctl[C_DONE]) return;
    const int it = ctl[C_ITERS], pi = (it - 1) & 1;
    auto *__restrict__ p = glw(d.p);
    auto *__restrict__ p32 = d.wp32 ? glw(d.p32) : nullptr;
    const auto *__restrict__ z = gl(d.z);
    const int64_t n = d.nnodes * 3, stride = (int64_t)d.nvblk * WGV;
    // the first dof's operands (and coarse term) before the partial sums
    int64_t i = (int64_t)ve.y * WGV + threadIdx.x;
    double pv = 0.0, zv = 0.0;
    if (i < n) {
        pv = p[i];
        zv = d.tl ? z[i] + coarse_term(d, i / 3, (int)(i % 3)) : z[i];
    }
    // rz of the previous iteration was published in the other parity slot;
    // the first workgroup publishes this one's and tests convergence
    const double rzn = sum_rz(d, pi);
    const double rzo = glw(d.scal)[rz_pub(1 - pi)];
    const double beta = rzn / rzo;
    if (ve.y == 0) {
        const double rr = sum_partials<WGV>(d.part + (pi ? P_RR1 : P_RR0), d.nvblk);
        if (threadIdx.x == 0) {
            glw(d.scal)[S_RR] = rr;
            glw(d.scal)[S_RZ] = rzn;
            glw(d.scal)[rz_pub(pi)] = rzn;
            if (rr <= glw(d.scal)[S_THR] || it >= d.maxiter || !std::isfinite(rr)) ctl[C_STOP] = 1;
        }
    }
    while (i < n) {
        const double pn = fma(beta, pv, zv);
        p[i] = pn;
        if (p32) p32[i] = (float)pn;
        i += stride;
        if (i < n) {
            pv = p[i];
            zv = d.tl ? z[i] + coarse_term(d, i / 3, (int)(i % 3)) : z[i];
        }
    }
}

// Two-level pieces, for the members with d.tl (mode 0: inside the iteration,
// skipped once the member is done; mode 1: initialisation and resume)
//   restrict: c_a = sum_{i in a} R_i^T r_i, one wave per aggregate
//   coarse:   y = Kc^-1 c, one wave per coarse row, and partials of c.y
//   prolong:  p = z + Phi y (start of a search direction)
__global__ void __launch_bounds__(64) tl_restrict_kernel(const SegDesc *__restrict__ D, const int2 *__restrict__ al,
                                                         int mode)
{
    const int2 e = al[blockIdx.x];
    if (e.x < 0) return;
    const SegDesc &d = D[e.x];
    if (mode == 0 && glw(d.ctl)[C_DONE]) return;
    const int32_t a = e.y;
    const auto *cent = gl(d.cent), *X = gl(d.X), *rv = gl(d.r);
    const auto *aggptr = gl(d.aggptr), *aggnodes = gl(d.aggnodes);
    const double cx = cent[a * 3], cy = cent[a * 3 + 1], cz = cent[a * 3 + 2];
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    // a lane's nodes k = k0 + lane + 64 j in order of j, eight at a time:
    // the eight nodes' loads are issued before their sums
    constexpr int NB = 8;
    const int32_t k1 = aggptr[a + 1];
    for (int32_t kb = aggptr[a] + threadIdx.x; kb < k1; kb += 64 * NB) {
        int64_t nd[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) nd[j] = kb + 64 * j < k1 ? aggnodes[kb + 64 * j] : -1;
        double r[NB][3], x[NB][3];
#pragma unroll
        for (int j = 0; j < NB; j++)
            if (nd[j] >= 0)
                for (int c = 0; c < 3; c++) {
                    r[j][c] = rv[nd[j] * 3 + c];
                    x[j][c] = X[nd[j] * 3 + c];
                }
#pragma unroll
        for (int j = 0; j < NB; j++)
            if (nd[j] >= 0) {
                const double r0 = r[j][0], r1 = r[j][1], r2 = r[j][2];
                const double dx = x[j][0] - cx, dy = x[j][1] - cy, dz = x[j][2] - cz;
                v[0] += r0;
                v[1] += r1;
                v[2] += r2;
                v[3] += fma(dy, r2, -dz * r1);
                v[4] += fma(dz, r0, -dx * r2);
                v[5] += fma(dx, r1, -dy * r0);
            }
    }
    for (int q = 0; q < 6; q++)
        for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; q++) glw(d.cvec)[a * 6 + q] = v[q];
}

__global__ void __launch_bounds__(256) tl_coarse_kernel(const SegDesc *__restrict__ D, const int2 *__restrict__ wl,
                                                        int mode)
{
    const int2 e = wl[blockIdx.x];
    if (e.x < 0) return;
    const SegDesc &d = D[e.x];
    if (mode == 0 && glw(d.ctl)[C_DONE]) return;
    const int pi = (glw(d.ctl)[C_ITERS] - 1) & 1;
    const int row = e.y * 4 + (int)(threadIdx.x / 64), lane = threadIdx.x % 64;
    double acc = 0.0;
    const auto *cvec = gl(d.cvec);
    if (row < d.nc) {
        // columns c = lane + 64 j in order of j, sixteen loads at a time
        const auto *Kr = gl(d.Kinv) + (int64_t)row * d.nc;
        constexpr int NB = 16;
        for (int cb = lane; cb < d.nc; cb += 64 * NB) {
            double kv[NB], cv[NB];
#pragma unroll
            for (int j = 0; j < NB; j++) {
                const bool in = cb + 64 * j < d.nc;
                kv[j] = in ? Kr[cb + 64 * j] : 0.0;
                cv[j] = in ? cvec[cb + 64 * j] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < NB; j++)
                if (cb + 64 * j < d.nc) acc = fma(kv[j], cv[j], acc);
        }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    __shared__ double cyw[4];
    if (lane == 0) {
        if (row < d.nc) glw(d.yvec)[row] = acc;
        cyw[threadIdx.x / 64] = row < d.nc ? cvec[row] * acc : 0.0;
    }
    __syncthreads();
    if (threadIdx.x == 0) glw(d.part)[(pi ? P_CY1 : P_CY0) + e.y] = ((cyw[0] + cyw[1]) + cyw[2]) + cyw[3];
}

__global__ void __launch_bounds__(WGV) tl_prolong_kernel(const SegDesc *__restrict__ D, const int2 *__restrict__ vl)
{
    const int2 ve = vl[blockIdx.x];
    const SegDesc &d = D[ve.x];
    float *__restrict__ p32 = d.wp32 ? d.p32 : nullptr;
    const int64_t n = d.nnodes * 3, stride = (int64_t)d.nvblk * WGV;
    for (int64_t i = (int64_t)ve.y * WGV + threadIdx.x; i < n; i += stride) {
        const double pn = d.z[i] + coarse_term(d, i / 3, (int)(i % 3));
        d.p[i] = pn;
        if (p32) p32[i] = (float)pn;
    }
}

// x = 0, r = b, z = M r, p = z; partials rz -> P_RZ1, rr -> P_RR1
__global__ void __launch_bounds__(WGV) pcg_init_kernel(const SegDesc *__restrict__ D, const int2 *__restrict__ vl)
{
    const int2 ve = vl[blockIdx.x];
    const SegDesc &d = D[ve.x];
    // iteration 0: the two-level kernels after this one write the parity
    // slot of C_ITERS - 1, i.e. slot 1, like the Jacobi partials below
    if (ve.y == 0 && threadIdx.x == 0) glw(d.ctl)[C_ITERS] = 0;
    double rz = 0.0, rr = 0.0;
    float *__restrict__ p32 = d.wp32 ? d.p32 : nullptr;
    const int64_t stride = (int64_t)d.nvblk * WGV;
    for (int64_t nd = (int64_t)ve.y * WGV + threadIdx.x; nd < d.nnodes; nd += stride) {
        double rv[3], zv[3];
        for (int c = 0; c < 3; c++) rv[c] = d.b[nd * 3 + c];
        apply_prec(d, nd, rv, zv);
        for (int c = 0; c < 3; c++) {
            const int64_t i = nd * 3 + c;
            d.x[i] = 0.0;
            d.r[i] = rv[c];
            d.z[i] = zv[c];
            d.p[i] = zv[c];
            if (p32) p32[i] = (float)zv[c];
            rz = fma(rv[c], zv[c], rz);
            rr = fma(rv[c], rv[c], rr);
        }
    }
    rz = block_sum<WGV>(rz);
    rr = block_sum<WGV>(rr);
    if (threadIdx.x == 0) {
        glw(d.part)[P_RZ1 + ve.y] = rz;
        glw(d.part)[P_RR1 + ve.y] = rr;
    }
}

// r = rt; z = M r; with restart also p = z (the search direction starts
// afresh); rz partials into the slot the next iteration reads as "previous".
__global__ void __launch_bounds__(WGV) pcg_rez_kernel(const SegDesc *__restrict__ D, const int2 *__restrict__ vl,
                                                      int restart)
{
    const int2 ve = vl[blockIdx.x];
    const SegDesc &d = D[ve.x];
    const int slot_pi = 1 - (glw(d.ctl)[C_ITERS] & 1);
    float *__restrict__ p32 = d.wp32 ? d.p32 : nullptr;
    double rz = 0.0;
    const int64_t stride = (int64_t)d.nvblk * WGV;
    for (int64_t nd = (int64_t)ve.y * WGV + threadIdx.x; nd < d.nnodes; nd += stride) {
        double rv[3], zv[3];
        for (int c = 0; c < 3; c++) rv[c] = d.rt[nd * 3 + c];
        apply_prec(d, nd, rv, zv);
        for (int c = 0; c < 3; c++) {
            const int64_t i = nd * 3 + c;
            d.r[i] = rv[c];
            d.z[i] = zv[c];
            if (restart) {
                d.p[i] = zv[c];
                if (p32) p32[i] = (float)zv[c];
            }
            rz = fma(rv[c], zv[c], rz);
        }
    }
    rz = block_sum<WGV>(rz);
    if (threadIdx.x == 0) glw(d.part)[(slot_pi ? P_RZ1 : P_RZ0) + ve.y] = rz;
}

// One workgroup per listed model publishes scalars from partials.
//   mode 0 (after init):        rz, rr = bb from slot 1; recursive threshold; reset control
//   mode 1 (after replacement): rr, bb, kx, drift from the residual partials
//   mode 2 (resume):            rz from the slot written by pcg_rez; clear the stop flags
__global__ void __launch_bounds__(WGV) pcg_finalize_kernel(const SegDesc *__restrict__ D,
                                                           const int32_t *__restrict__ sl, int mode)
{
    const SegDesc &d = D[sl[blockIdx.x]];
    auto *ctl = glw(d.ctl);
    double *scal = d.scal;
    if (mode == 0) {
        const double rz = sum_rz(d, 1);
        const double rr = sum_partials<WGV>(d.part + P_RR1, d.nvblk);
        if (threadIdx.x == 0) {
            scal[S_RZ] = rz;
            scal[rz_pub(1)] = rz;
            scal[S_RR] = rr;
            scal[S_BB] = rr;
            scal[S_KX] = 0.0;
            scal[S_GAP] = 0.0;
            scal[S_THR] = d.thr_fac * rr;
            ctl[C_ITERS] = 0;
            ctl[C_BREAK] = 0;
            ctl[C_STOP] = 0;
            ctl[C_DONE] = rr == 0.0 ? 1 : 0;
        }
    } else if (mode == 1) {
        const double rr = sum_partials<WGV>(d.part + P_A, d.ntiles);
        const double bb = sum_partials<WGV>(d.part + P_B, d.ntiles);
        const double kx = sum_partials<WGV>(d.part + P_C, d.ntiles);
        const double gap = sum_partials<WGV>(d.part + P_D, d.ntiles);
        if (threadIdx.x == 0) {
            scal[S_RR] = rr;
            scal[S_BB] = bb;
            scal[S_KX] = kx;
            scal[S_GAP] = gap;
        }
    } else {
        const int slot_pi = 1 - (ctl[C_ITERS] & 1);
        const double rz = sum_rz(d, slot_pi);
        if (threadIdx.x == 0) {
            scal[S_RZ] = rz;
            scal[rz_pub(slot_pi)] = rz;
            ctl[C_DONE] = 0;
            ctl[C_STOP] = 0;
            ctl[C_BREAK] = 0;
        }
    }
}

__global__ void fill_f32_kernel(int64_t n, float *__restrict__ x, float v)
{
    for (int64_t i = (int64_t)blockIdx.x * WGV + threadIdx.x; i < n; i += (int64_t)gridDim.x * WGV) x[i] = v;
}

__global__ void fill_kernel(int64_t n, double *__restrict__ x, double v)
{
    for (int64_t i = (int64_t)blockIdx.x * WGV + threadIdx.x; i < n; i += (int64_t)gridDim.x * WGV) x[i] = v;
}

// Row tiles of a model: R = 64 rows, a whole number of the row passes of a
// 512-thread workgroup at every lanes value (16 rows per pass at 32 lanes,
// 8 at 64, 32 at 16, 64 at 8); models of more than 64 * MAX_GR nodes take
// the next multiple of 64 that keeps MAX_GR tiles.  On the 40^3 HEX20 block
// (L = 32) the SpMV took 153.9 us with R = 64 and 154.5 us with 48 against
// 155.8-160.0 us with 40, 56, 72, 80 and 96 (rows that leave the last pass
// partly idle, or longer tiles) and 162.6 us with 128.  Vector kernels: at
// most MAX_GV workgroups.
struct Geom {
    int R, ntiles, nvblk;
};

Geom model_geom(const amr_model *m)
{
    Geom g;
    const int64_t n = std::max<int64_t>(1, m->nnodes);
    int64_t R = 64;
    if ((n + R - 1) / R > MAX_GR) R = ((n + MAX_GR - 1) / MAX_GR + 63) / 64 * 64;
    g.R = (int)R;
    g.ntiles = (int)((n + R - 1) / R);
    const int64_t nv = (m->nnodes * 3 + WGV - 1) / WGV;
    g.nvblk = (int)std::max<int64_t>(1, std::min<int64_t>(nv, MAX_GV));
    return g;
}

size_t tile_lds(int R, int ndbl) { return (size_t)(((R + 1) * 4 + 15) & ~15) + (size_t)ndbl * pow2ceil(R) * 8; }

// SpMV operand kinds: FP64 values; FP32 values with FP32 p and FP32
// accumulation; FP32 values with FP64 gather and accumulation.
enum { K_F64 = 0, K_F32, K_F32D, K_NKIND };

int lanes_index(int lanes) { return lanes == 8 ? 0 : lanes == 16 ? 1 : lanes == 32 ? 2 : 3; }
constexpr int LANES[4] = {8, 16, 32, 64};

int kind_lanes(const amr_model *m, int kind)
{
    return kind == K_F64 ? m->lanes64 : kind == K_F32 ? m->lanes32 : m->lanes32d;
}

void launch_pcg_spmv(int kind, int lanes, unsigned grid, int Rmax, hipStream_t s, const SegDesc *D, const int2 *tl)
{
    const size_t lds = tile_lds(Rmax, 1);
#define AMR_PCG_SPMV(LN, VT, XT, AT) \
    hipLaunchKernelGGL((pcg_spmv_kernel<LN, VT, XT, AT>), dim3(grid), dim3(WGS), lds, s, D, tl)
#define AMR_PCG_KIND(LN)                                          \
    do {                                                          \
        if (kind == K_F64) AMR_PCG_SPMV(LN, double, double, double); \
        else if (kind == K_F32) AMR_PCG_SPMV(LN, float, float, float); \
        else AMR_PCG_SPMV(LN, float, double, double);             \
    } while (0)
    switch (lanes) {
    case 8: AMR_PCG_KIND(8); break;
    case 16: AMR_PCG_KIND(16); break;
    case 32: AMR_PCG_KIND(32); break;
    default: AMR_PCG_KIND(64); break;
    }
#undef AMR_PCG_KIND
#undef AMR_PCG_SPMV
}

void launch_residual(int lanes, unsigned grid, int Rmax, hipStream_t s, const SegDesc *D, const int2 *tl)
{
    const size_t lds = tile_lds(Rmax, 4);
#define AMR_RES(LN) hipLaunchKernelGGL((residual_kernel<LN>), dim3(grid), dim3(WGS), lds, s, D, tl)
    switch (lanes) {
    case 8: AMR_RES(8); break;
    case 16: AMR_RES(16); break;
    case 32: AMR_RES(32); break;
    default: AMR_RES(64); break;
    }
#undef AMR_RES
}

}  // namespace

// ------------------------------------------------------------- the pool

namespace {

struct Member {
    amr_model *m = nullptr;
    amr_solve_opts o{};
    const double *b = nullptr;
    double *x = nullptr;
    int state = 0;  // 0 free, 1 joining, 2 active
    bool mixed = false, f32x = false;
    double bb = 0.0, rr_ref = 0.0, rr_true_prev = INFINITY, relres = 0.0;
    int converged = 0, broke = 0, replacements = 0, iters = 0, fell_back = 0, breakdowns = 0;
    int last_rep = 0;  // iteration of the last residual replacement
    bool form = false;  // its coarse operator is to be formed before it starts
    // per-step scratch
    bool rep = false, done = false;
    double rr_rec = 0.0;
};

// M_WAIT: joined, but work the model queued on its own stream before
// joining (assembly), or the formation of its coarse operator, may still be
// running; it starts at the first step that finds that work done, so that a
// slow joiner does not hold up the members already solving
enum { M_FREE = 0, M_JOIN, M_ACTIVE, M_WAIT };

}  // namespace

struct amr_pool {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int cap = 0, check = 32;
    std::vector<Member> mem;
    std::vector<amr_model *> attached;
    // device descriptors and scalars, with pinned host mirrors
    SegDesc *dD = nullptr, *hD = nullptr;
    double *dS = nullptr, *hS = nullptr;
    // work lists: staged in pinned memory, uploaded in one copy per use
    char *dL = nullptr, *hL = nullptr;
    size_t lcap = 0, lused = 0;
    bool pending = false;  // an upload from the pinned mirrors may still be in flight
    hipEvent_t ev_upload = nullptr;  // recorded after each upload
    bool desc_dirty = false;
    // a step launched by amr_pool_step_begin and not yet collected
    bool in_flight = false;
    hipEvent_t ev_step = nullptr;    // recorded after the step's scalar read-back
    std::vector<int> fl_act;         // members of the in-flight step
    // models that left the pool at the last step and still run on its
    // stream; they return to their own streams when the next step is queued
    // (or at amr_pool_hand_back), so that the work queued for them in
    // between (the NR finish) runs ahead of that step
    std::vector<amr_model *> leaving;
    // SpMV timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    double t_spmv_ms = 0.0;
    int64_t n_spmv = 0;
    // captured iteration blocks of a step (check_every iterations of every
    // PCG phase), keyed by their launch shapes; replayed while the shapes
    // stay (work lists padded to size classes, members joining and leaving
    // change only the lists' contents)
    struct StepGraph {
        std::vector<int64_t> key;
        hipGraphExec_t exec = nullptr;
        uint64_t used = 0;
    };
    std::vector<StepGraph> graphs;
    bool use_graphs = false;
    uint64_t graph_clock = 0;
    int64_t graph_captures = 0, graph_replays = 0;
    // coarse operators of two-level joiners are formed in batches: on the
    // side stream for joiners that start at a later step (the members
    // already solving are not held up), on the pool's stream for those that
    // start at once
    hipStream_t side = nullptr;
    AmrStaging stg_side, stg_pool;
    int64_t coarse_batches = 0, coarse_formed = 0;
    int64_t iter_kernels = 0, iter_submits = 0;  // of the steps' PCG iterations
    int64_t member_kernels = 0;  // the same iterations launched member by member
};

namespace {

int pool_sync(amr_pool *p)
{
    AMR_HIP(hipStreamSynchronize(p->stream));
    p->pending = false;
    return AMR_OK;
}

// Make the pinned mirrors writable: wait for any upload still reading them
// (only for the copy, not for the kernels queued after it).
int pool_begin_host_writes(amr_pool *p)
{
    if (!p->pending) return AMR_OK;
    AMR_HIP(hipEventSynchronize(p->ev_upload));
    p->pending = false;
    return AMR_OK;
}

void drop_graphs(amr_pool *p)
{
    for (auto &g : p->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    p->graphs.clear();
}

int pool_reserve_lists(amr_pool *p, size_t bytes)
{
    if (bytes <= p->lcap) return AMR_OK;
    // kernels queued earlier may still read the device list buffer
    if (int rc = pool_sync(p)) return rc;
    drop_graphs(p);  // they point into the old buffer
    size_t ncap = std::max<size_t>(bytes, 2 * p->lcap);
    ncap = std::max<size_t>(ncap, 1 << 16);
    if (p->dL) AMR_HIP(hipFree(p->dL));
    if (p->hL) AMR_HIP(hipHostFree(p->hL));
    p->dL = p->hL = nullptr;
    p->lcap = 0;
    AMR_HIP(hipMalloc(&p->dL, ncap));
    AMR_HIP(hipHostMalloc(&p->hL, ncap, hipHostMallocDefault));
    p->lcap = ncap;
    return AMR_OK;
}

// List staging: reset, append, upload.  Offsets are 16-byte aligned.
struct Lists {
    amr_pool *p;
    size_t off = 0;
    template <typename T>
    T *dev_at(size_t o) const { return (T *)(p->dL + o); }
    template <typename T>
    size_t push(const std::vector<T> &v)
    {
        const size_t o = off;
        if (!v.empty()) std::memcpy(p->hL + o, v.data(), v.size() * sizeof(T));
        off = (o + v.size() * sizeof(T) + 15) & ~(size_t)15;
        return o;
    }
};

template <typename T>
size_t padded(const std::vector<T> &v) { return (v.size() * sizeof(T) + 15) & ~(size_t)15; }

int upload(amr_pool *p, size_t bytes, bool desc)
{
    if (bytes) AMR_HIP(hipMemcpyAsync(p->dL, p->hL, bytes, hipMemcpyHostToDevice, p->stream));
    if (desc) {
        AMR_HIP(hipMemcpyAsync(p->dD, p->hD, sizeof(SegDesc) * p->cap, hipMemcpyHostToDevice, p->stream));
        p->desc_dirty = false;
    }
    if (bytes || desc) {
        AMR_HIP(hipEventRecord(p->ev_upload, p->stream));
        p->pending = true;
    }
    return AMR_OK;
}

// A model's work runs on its own stream, and on the pool's stream while it
// is a member; the switches are ordered with events, without waiting.
int handoff(amr_model *m, hipStream_t from, hipStream_t to)
{
    if (from == to) return AMR_OK;
    if (!m->ev_handoff) AMR_HIP(hipEventCreateWithFlags(&m->ev_handoff, hipEventDisableTiming));
    AMR_HIP(hipEventRecord(m->ev_handoff, from));
    AMR_HIP(hipStreamWaitEvent(to, m->ev_handoff, 0));
    m->stream = to;
    return AMR_OK;
}

void fill_desc(amr_pool *p, int slot)
{
    Member &M = p->mem[slot];
    amr_model *m = M.m;
    const Geom g = model_geom(m);
    SegDesc &d = p->hD[slot];
    std::memset(&d, 0, sizeof d);
    d.rowptr = m->rowptr;
    d.col = m->col;
    d.col16 = m->col16;
    d.Kv = m->Kv;
    d.Kv32 = m->Kv32;
    d.mask = m->mask;
    d.dinv = m->dinv;
    d.binv = m->binv;
    d.dinv32 = m->dinv32;
    d.binv32 = m->binv32;
    d.b = M.b;
    d.x = M.x;
    d.r = m->v[1];
    d.z = m->v[2];
    d.p = m->v[3];
    d.q = m->v[4];
    d.rt = m->v[5];
    d.p32 = m->f[0];
    d.part = m->part;
    d.scal = p->dS + (size_t)slot * SEG_DBL;
    d.ctl = (int32_t *)(d.scal + S_NSCAL);
    d.nnodes = m->nnodes;
    d.nblk = m->pat.nblk;
    const double tr = 0.5 * M.o.rtol;
    d.thr_fac = tr * tr;
    d.R = g.R;
    d.ntiles = g.ntiles;
    d.nvblk = g.nvblk;
    d.maxiter = M.o.maxiter;
    d.f32 = M.mixed ? 1 : 0;
    d.block = M.o.precond == AMR_PRECOND_BLOCK_JACOBI ? 1 : 0;
    d.wp32 = (M.mixed && M.f32x) ? 1 : 0;
    d.tl = M.o.precond == AMR_PRECOND_TWO_LEVEL ? 1 : 0;
    if (d.tl) {
        d.nagg = m->nagg;
        d.nc = m->ncoarse;
        d.ncw = (m->ncoarse + 3) / 4;
        d.agg = m->c_agg;
        d.aggptr = m->c_aggptr;
        d.aggnodes = m->c_aggnodes;
        d.cent = m->c_cent;
        d.X = m->X;
        d.Kinv = m->Kinv;
        d.cvec = m->c_vec;
        d.yvec = m->y_vec;
    }
}

std::vector<int2> agg_list(amr_pool *p, const std::vector<int> &slots)
{
    std::vector<int2> v;
    for (int s : slots)
        if (p->hD[s].tl)
            for (int a = 0; a < p->hD[s].nagg; a++) v.push_back(make_int2(s, a));
    return v;
}

std::vector<int2> cw_list(amr_pool *p, const std::vector<int> &slots)
{
    std::vector<int2> v;
    for (int s : slots)
        if (p->hD[s].tl)
            for (int w = 0; w < p->hD[s].ncw; w++) v.push_back(make_int2(s, w));
    return v;
}

std::vector<int> tl_only(amr_pool *p, const std::vector<int> &slots)
{
    std::vector<int> v;
    for (int s : slots)
        if (p->hD[s].tl) v.push_back(s);
    return v;
}

// the two-level steps after an initialisation or a resume: c, y, and with
// `prolong` the search direction p = z + Phi y
void launch_tl_fresh(amr_pool *p, const Lists &Ls, size_t o_al, unsigned nal, size_t o_cw, unsigned ncw,
                     size_t o_pv, unsigned npv)
{
    hipStream_t st = p->stream;
    if (nal) hipLaunchKernelGGL(tl_restrict_kernel, dim3(nal), dim3(64), 0, st, p->dD, Ls.dev_at<int2>(o_al), 1);
    if (ncw) hipLaunchKernelGGL(tl_coarse_kernel, dim3(ncw), dim3(256), 0, st, p->dD, Ls.dev_at<int2>(o_cw), 1);
    if (npv) hipLaunchKernelGGL(tl_prolong_kernel, dim3(npv), dim3(WGV), 0, st, p->dD, Ls.dev_at<int2>(o_pv));
}

int member_kind(const Member &M) { return !M.mixed ? K_F64 : (M.f32x ? K_F32 : K_F32D); }

// Tile lists of the given members grouped by (kind, lanes); for residual
// launches kind is ignored and the FP64 lanes are used.
struct Group {
    int kind, lanes, Rmax;
    size_t off;     // device offset of the group's entries
    unsigned n;
};

void build_groups(amr_pool *p, const std::vector<int> &slots, bool residual, std::vector<std::vector<int2>> &ent,
                  std::vector<Group> &gr)
{
    ent.assign(K_NKIND * 4, {});
    std::vector<int> rmax(K_NKIND * 4, 0);
    for (int s : slots) {
        const Member &M = p->mem[s];
        const int kind = residual ? K_F64 : member_kind(M);
        const int key = kind * 4 + lanes_index(kind_lanes(M.m, kind));
        const SegDesc &d = p->hD[s];
        for (int t = 0; t < d.ntiles; t++) ent[key].push_back(make_int2(s, t));
        rmax[key] = std::max(rmax[key], d.R);
    }
    gr.clear();
    for (int key = 0; key < K_NKIND * 4; key++)
        if (!ent[key].empty()) gr.push_back({key / 4, LANES[key % 4], rmax[key], 0, (unsigned)ent[key].size()});
}

// Work lists of a captured step are padded with empty entries (member -1,
// which the kernels skip) to a size class: 64, then 2^k and 3*2^(k-1).
size_t size_class(size_t n)
{
    if (n <= 64) return 64;
    size_t c = 64;
    while (c < n) {
        if (c + c / 2 >= n) return c + c / 2;
        c *= 2;
    }
    return c;
}

void pad_list(std::vector<int2> &v)
{
    if (!v.empty()) v.resize(size_class(v.size()), make_int2(-1, -1));
}

std::vector<int2> vec_list(amr_pool *p, const std::vector<int> &slots)
{
    std::vector<int2> v;
    for (int s : slots)
        for (int j = 0; j < p->hD[s].nvblk; j++) v.push_back(make_int2(s, j));
    return v;
}

int ensure_events(amr_pool *p)
{
    if ((int)p->ev.size() >= 2 * p->check) return AMR_OK;
    for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
    p->ev.assign(2 * p->check, nullptr);
    for (hipEvent_t &e : p->ev) AMR_HIP(hipEventCreate(&e));
    return AMR_OK;
}

void finish(amr_pool *p, int slot, std::vector<amr_pool_result> &out)
{
    Member &M = p->mem[slot];
    amr_pool_result r;
    r.m = M.m;
    r.info.iterations = M.iters;
    r.info.refinements = M.replacements;
    r.info.converged = M.converged;
    r.info.fallback = M.fell_back;
    r.info.relres = M.relres;
    out.push_back(r);
    M = Member();
}

}  // namespace

int amr_pool_new(amr_pool **out, int device, int cap, int check, hipStream_t stream)
{
    if (cap < 1 || cap > (1 << 16)) {
        amr_set_error("amr_pool_create: max_members must be in [1, 65536]");
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipSetDevice(device));
    amr_pool *p = new amr_pool();
    p->device = device;
    p->cap = cap;
    p->check = std::max(1, std::min(check > 0 ? check : 32, 256));
    p->mem.assign(cap, Member());
    // steps launch kernel by kernel unless AMR_POOL_GRAPHS=1 (or
    // amr_pool_set_graphs) asks for captured ones: on MI355X the replayed
    // graphs ran the pooled workloads 4-12% slower than the same kernels
    // launched one by one
    if (const char *g = std::getenv("AMR_POOL_GRAPHS")) p->use_graphs = std::atoi(g) != 0;
    int rc = AMR_OK;
    auto fail = [&](hipError_t e) {
        amr_set_error(std::string("HIP error: ") + hipGetErrorString(e) + " creating a pool");
        amr_pool_free(p);
        return AMR_ERR_HIP;
    };
    hipError_t e;
    if (stream) p->stream = stream;
    else {
        if ((e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e);
        p->own_stream = true;
    }
    if ((e = hipEventCreateWithFlags(&p->ev_upload, hipEventDisableTiming)) != hipSuccess) return fail(e);
    if ((e = hipEventCreateWithFlags(&p->ev_step, hipEventDisableTiming)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&p->dD, sizeof(SegDesc) * cap)) != hipSuccess) return fail(e);
    if ((e = hipHostMalloc(&p->hD, sizeof(SegDesc) * cap, hipHostMallocDefault)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&p->dS, sizeof(double) * SEG_DBL * cap)) != hipSuccess) return fail(e);
    if ((e = hipHostMalloc(&p->hS, sizeof(double) * SEG_DBL * cap, hipHostMallocDefault)) != hipSuccess)
        return fail(e);
    std::memset(p->hD, 0, sizeof(SegDesc) * cap);
    if ((e = hipMemsetAsync(p->dS, 0, sizeof(double) * SEG_DBL * cap, p->stream)) != hipSuccess) return fail(e);
    if ((e = hipStreamSynchronize(p->stream)) != hipSuccess) return fail(e);
    (void)rc;
    *out = p;
    return AMR_OK;
}

void amr_pool_free(amr_pool *p)
{
    if (!p) return;
    (void)hipSetDevice(p->device);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    for (amr_model *m : p->attached) {
        m->stream = m->own_stream;
        m->pool = nullptr;
    }
    if (p->dD) (void)hipFree(p->dD);
    if (p->hD) (void)hipHostFree(p->hD);
    if (p->dS) (void)hipFree(p->dS);
    if (p->hS) (void)hipHostFree(p->hS);
    if (p->dL) (void)hipFree(p->dL);
    if (p->hL) (void)hipHostFree(p->hL);
    for (hipEvent_t e : p->ev)
        if (e) (void)hipEventDestroy(e);
    if (p->ev_upload) (void)hipEventDestroy(p->ev_upload);
    if (p->ev_step) (void)hipEventDestroy(p->ev_step);
    if (p->side) (void)hipStreamSynchronize(p->side);
    amr_staging_free(p->stg_side);
    amr_staging_free(p->stg_pool);
    if (p->side) (void)hipStreamDestroy(p->side);
    drop_graphs(p);
    if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

int amr_pool_attach_model(amr_pool *p, amr_model *m)
{
    if (m->pool == p) return AMR_OK;
    if (m->pool) {
        amr_set_error("the model is attached to another pool");
        return AMR_ERR_ARG;
    }
    if (m->device != p->device) {
        amr_set_error("the model and the pool are on different devices");
        return AMR_ERR_ARG;
    }
    m->pool = p;
    p->attached.push_back(m);
    return AMR_OK;
}

int amr_pool_detach_model(amr_model *m)
{
    amr_pool *p = m->pool;
    if (!p) return AMR_OK;
    bool member = false;
    for (int s = 0; s < p->cap; s++)
        if (p->mem[s].m == m) {
            p->mem[s] = Member();
            member = true;
        }
    if (member || m->stream != m->own_stream) {
        // its kernels may still be queued on the pool's stream (or, for its
        // coarse operator, on the side stream)
        AMR_HIP(hipStreamSynchronize(p->stream));
        if (p->side) AMR_HIP(hipStreamSynchronize(p->side));
        p->fl_act.erase(std::remove_if(p->fl_act.begin(), p->fl_act.end(),
                                       [&](int s) { return p->mem[s].state == M_FREE; }),
                        p->fl_act.end());
        m->stream = m->own_stream;
    }
    p->attached.erase(std::remove(p->attached.begin(), p->attached.end(), m), p->attached.end());
    p->leaving.erase(std::remove(p->leaving.begin(), p->leaving.end(), m), p->leaving.end());
    m->pool = nullptr;
    return AMR_OK;
}

bool amr_pool_is_member(const amr_model *m)
{
    if (!m->pool) return false;
    for (const Member &M : m->pool->mem)
        if (M.m == m) return true;
    return false;
}

int amr_pool_count(const amr_pool *p)
{
    int n = 0;
    for (const Member &M : p->mem) n += M.state != M_FREE;
    return n;
}

// Join: the model's system K11 x = b starts at the next step.
int amr_pool_join(amr_pool *p, amr_model *m, const double *b, double *x, const amr_solve_opts *o)
{
    if (m->pool != p && m->self_pool != p) {
        amr_set_error("amr_pool_add: the model is not attached to this pool");
        return AMR_ERR_ARG;
    }
    for (const Member &M : p->mem)
        if (M.m == m) {
            amr_set_error("amr_pool_add: the model is already solving in this pool");
            return AMR_ERR_ARG;
        }
    int slot = -1;
    for (int s = 0; s < p->cap && slot < 0; s++)
        if (p->mem[s].state == M_FREE) slot = s;
    if (slot < 0) {
        amr_set_error("amr_pool_add: the pool is full");
        return AMR_ERR_ARG;
    }
    if (int rc = amr_choose_lanes(m)) return rc;
    bool form = false;
    if (o->precond == AMR_PRECOND_TWO_LEVEL) {
        if (int rc = amr_coarse_build(m)) return rc;
        if ((m->ncoarse + 3) / 4 > MAX_CW) {
            amr_set_error("two-level preconditioner: too many aggregates for the coarse product");
            return AMR_ERR_ARG;
        }
        form = amr_coarse_due(m);
    }
    Member M;
    M.form = form;
    M.m = m;
    M.o = *o;
    M.b = b;
    M.x = x;
    M.state = M_JOIN;
    M.mixed = o->precision == AMR_PREC_MIXED || o->precision == AMR_PREC_MIXED_F64ACC;
    M.f32x = o->precision == AMR_PREC_MIXED;
    if (int rc = pool_begin_host_writes(p)) return rc;
    // the model's queued work (assembly, right-hand side) precedes its solve
    if (m->stream != p->stream) {
        if (!m->ev_handoff) AMR_HIP(hipEventCreateWithFlags(&m->ev_handoff, hipEventDisableTiming));
        AMR_HIP(hipEventRecord(m->ev_handoff, m->stream));
        m->stream = p->stream;
        M.state = M_WAIT;
    }
    p->mem[slot] = M;
    fill_desc(p, slot);
    p->desc_dirty = true;
    return AMR_OK;
}

// A step: initialise joining members, run check_every PCG iterations for
// every member, then residual replacements and convergence decisions.
// amr_pool_step_launch queues the iterations and the read-back of the
// members' scalars without waiting; amr_pool_step_collect waits for them,
// decides, queues the replacements and resumes, and appends the members
// that finished (converged or failed) to `out`; they leave the pool, and
// their models' work returns to their own streams when the next step is
// queued.  Members that join while a step is in flight start at the next
// one.
int amr_pool_step_launch(amr_pool *p)
{
    if (p->in_flight) {
        amr_set_error("amr_pool_step_begin: the previous step has not been collected");
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipSetDevice(p->device));
    // joiners whose own work is done start now; if nobody else is solving,
    // the waiting ones start too, the pool's stream waiting for them
    // (a joiner whose coarse operator is due is formed first: on the side
    // stream when others can start now, else on the pool's stream)
    std::vector<int> waiting, side;
    bool ready = false;
    for (int s = 0; s < p->cap; s++) {
        Member &M = p->mem[s];
        if (M.state == M_WAIT) {
            const hipError_t q = hipEventQuery(M.m->ev_handoff);
            if (q == hipSuccess) {
                if (M.form) side.push_back(s);
                else M.state = M_JOIN;
            } else if (q == hipErrorNotReady) {
                (void)hipGetLastError();  // not an error: clear it
                waiting.push_back(s);
            } else {
                AMR_HIP(q);
            }
        }
        ready = ready || M.state == M_JOIN || M.state == M_ACTIVE;
    }
    if (!ready) {
        for (int s : waiting) {
            AMR_HIP(hipStreamWaitEvent(p->stream, p->mem[s].m->ev_handoff, 0));
            p->mem[s].state = M_JOIN;
        }
        for (int s : side) p->mem[s].state = M_JOIN;
        side.clear();
    }
    if (!side.empty()) {
        if (!p->side) AMR_HIP(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
        std::vector<amr_model *> ms;
        for (int s : side) ms.push_back(p->mem[s].m);
        if (int rc = amr_coarse_form(p->side, ms.data(), (int)ms.size(), p->stg_side)) return rc;
        p->coarse_batches++;
        p->coarse_formed += (int64_t)ms.size();
        // they start at the first step that finds the formation done
        for (int s : side) {
            AMR_HIP(hipEventRecord(p->mem[s].m->ev_handoff, p->side));
            p->mem[s].form = false;
        }
    }
    {
        std::vector<amr_model *> ms;
        for (int s = 0; s < p->cap; s++)
            if (p->mem[s].state == M_JOIN && p->mem[s].form) {
                ms.push_back(p->mem[s].m);
                p->mem[s].form = false;
            }
        if (!ms.empty()) {
            if (int rc = amr_coarse_form(p->stream, ms.data(), (int)ms.size(), p->stg_pool)) return rc;
            p->coarse_batches++;
            p->coarse_formed += (int64_t)ms.size();
        }
    }
    std::vector<int> join, act;
    for (int s = 0; s < p->cap; s++) {
        if (p->mem[s].state == M_JOIN) join.push_back(s);
        if (p->mem[s].state == M_JOIN || p->mem[s].state == M_ACTIVE) act.push_back(s);
    }
    if (int rc = amr_pool_hand_back(p)) return rc;
    if (act.empty()) return AMR_OK;
    const int check = p->check;
    hipStream_t st = p->stream;

    // (a) lists for the iterations (first, so that their offsets depend on
    // their own sizes only) and for the initialisation of joiners
    const bool graph = p->use_graphs && !p->timing;
    std::vector<std::vector<int2>> ent;
    std::vector<Group> gr;
    if (int rc = pool_begin_host_writes(p)) return rc;
    build_groups(p, act, false, ent, gr);
    std::vector<int2> va = vec_list(p, act), aa = agg_list(p, act), ca = cw_list(p, act);
    if (graph) {
        for (auto &e : ent) pad_list(e);
        int gi = 0;
        for (const auto &e : ent)
            if (!e.empty()) {
                gr[gi].n = (unsigned)e.size();
                gr[gi].Rmax = (gr[gi].Rmax + 63) / 64 * 64;
                gi++;
            }
        pad_list(va);
        pad_list(aa);
        pad_list(ca);
    }
    std::vector<int2> vj = vec_list(p, join);
    std::vector<int32_t> sj(join.begin(), join.end());
    std::vector<int2> aj = agg_list(p, join), cj = cw_list(p, join), pj = vec_list(p, tl_only(p, join));
    size_t need = padded(vj) + padded(va) + padded(sj) + padded(aj) + padded(cj) + padded(pj) + padded(aa) +
                  padded(ca);
    for (const auto &e : ent) need += padded(e);
    if (int rc = pool_reserve_lists(p, need)) return rc;
    Lists Ls{p};
    const size_t o_va = Ls.push(va), o_aa = Ls.push(aa), o_ca = Ls.push(ca);
    {
        int gi = 0;
        for (const auto &e : ent)
            if (!e.empty()) gr[gi++].off = Ls.push(e);
    }
    const size_t o_vj = Ls.push(vj), o_sj = Ls.push(sj);
    const size_t o_aj = Ls.push(aj), o_cj = Ls.push(cj), o_pj = Ls.push(pj);
    if (int rc = upload(p, Ls.off, p->desc_dirty)) return rc;
    if (!join.empty()) {
        hipLaunchKernelGGL(pcg_init_kernel, dim3((unsigned)vj.size()), dim3(WGV), 0, st, p->dD,
                           Ls.dev_at<int2>(o_vj));
        launch_tl_fresh(p, Ls, o_aj, (unsigned)aj.size(), o_cj, (unsigned)cj.size(), o_pj, (unsigned)pj.size());
        hipLaunchKernelGGL(pcg_finalize_kernel, dim3((unsigned)sj.size()), dim3(WGV), 0, st, p->dD,
                           Ls.dev_at<int32_t>(o_sj), 0);
    }
    if (p->timing)
        if (int rc = ensure_events(p)) return rc;
    // (b) check_every iterations of every phase
    auto iterations = [&]() {
        for (int j = 0; j < check; j++) {
            if (p->timing) (void)hipEventRecord(p->ev[2 * j], st);
            for (const Group &g : gr)
                launch_pcg_spmv(g.kind, g.lanes, g.n, g.Rmax, st, p->dD, Ls.dev_at<int2>(g.off));
            if (p->timing) (void)hipEventRecord(p->ev[2 * j + 1], st);
            hipLaunchKernelGGL(pcg_update_kernel, dim3((unsigned)va.size()), dim3(WGV), 0, st, p->dD,
                               Ls.dev_at<int2>(o_va));
            if (!aa.empty()) {
                hipLaunchKernelGGL(tl_restrict_kernel, dim3((unsigned)aa.size()), dim3(64), 0, st, p->dD,
                                   Ls.dev_at<int2>(o_aa), 0);
                hipLaunchKernelGGL(tl_coarse_kernel, dim3((unsigned)ca.size()), dim3(256), 0, st, p->dD,
                                   Ls.dev_at<int2>(o_ca), 0);
            }
            hipLaunchKernelGGL(pcg_dir_kernel, dim3((unsigned)va.size()), dim3(WGV), 0, st, p->dD,
                               Ls.dev_at<int2>(o_va));
        }
    };
    // kernels of the step's iterations, and the host submissions that queue
    // them: one each kernel by kernel, one graph launch when captured
    const int64_t iter_kernels = (int64_t)check * ((int64_t)gr.size() + 2 + (aa.empty() ? 0 : 2));
    p->iter_kernels += iter_kernels;
    p->iter_submits += graph ? 1 : iter_kernels;
    // launching each member alone would take its own SpMV, update and
    // direction kernels (and restrict and coarse product, two-level) per
    // iteration
    for (int s : act) p->member_kernels += (int64_t)check * (p->hD[s].tl ? 5 : 3);
    if (!graph) {
        iterations();
    } else {
        std::vector<int64_t> key = {check, (int64_t)(intptr_t)p->dL, (int64_t)va.size(), (int64_t)aa.size(),
                                    (int64_t)ca.size()};
        for (const Group &g : gr) key.insert(key.end(), {g.kind, g.lanes, g.Rmax, (int64_t)g.n});
        amr_pool::StepGraph *hit = nullptr;
        for (auto &g : p->graphs)
            if (g.key == key) hit = &g;
        if (!hit) {
            // capture the iterations once for these shapes (no memset nodes:
            // the kernels write every value they later read)
            AMR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
            iterations();
            const hipError_t le = hipGetLastError();
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(st, &g);
            AMR_HIP(le);
            AMR_HIP(ce);
            hipGraphExec_t ex = nullptr;
            const hipError_t ie = hipGraphInstantiateWithFlags(&ex, g, 0);
            (void)hipGraphDestroy(g);
            AMR_HIP(ie);
            if (p->graphs.size() >= 8) {
                auto lru = std::min_element(p->graphs.begin(), p->graphs.end(),
                                            [](const auto &a, const auto &b) { return a.used < b.used; });
                AMR_HIP(hipStreamSynchronize(st));   // it may still be running
                (void)hipGraphExecDestroy(lru->exec);
                p->graphs.erase(lru);
            }
            p->graphs.push_back({key, ex, 0});
            hit = &p->graphs.back();
            p->graph_captures++;
        } else {
            p->graph_replays++;
        }
        hit->used = ++p->graph_clock;
        AMR_HIP(hipGraphLaunch(hit->exec, st));
    }
    AMR_HIP(hipGetLastError());
    const size_t sbytes = sizeof(double) * SEG_DBL * p->cap;
    AMR_HIP(hipMemcpyAsync(p->hS, p->dS, sbytes, hipMemcpyDeviceToHost, st));
    AMR_HIP(hipEventRecord(p->ev_step, st));
    p->fl_act = act;
    p->in_flight = true;
    return AMR_OK;
}

namespace {

int step_decide(amr_pool *p, std::vector<amr_pool_result> &out)
{
    const std::vector<int> act = p->fl_act;
    const int check = p->check;
    const double kEps = 2.220446049250313e-16, kFloorC = 64.0, kGrow = 1e4;
    const int kForceEvery = 256;
    hipStream_t st = p->stream;
    const size_t sbytes = sizeof(double) * SEG_DBL * p->cap;
    std::vector<std::vector<int2>> ent;
    std::vector<Group> gr;
    Lists Ls{p};
    size_t need = 0;

    // (b) decisions; residual replacement for members that stopped or whose
    // recursive residual dropped enough (mixed paths)
    int real_max = 0;
    std::vector<int> rep;
    for (int s : act) {
        Member &M = p->mem[s];
        const double *hs = p->hS + (size_t)s * SEG_DBL;
        const int32_t *hc = (const int32_t *)(hs + S_NSCAL);
        const int prev = M.state == M_JOIN ? 0 : M.iters;
        if (M.state == M_JOIN) {
            M.state = M_ACTIVE;
            M.bb = hs[S_BB];
            M.rr_ref = M.bb;
            if (M.bb == 0.0) {
                M.converged = 1;
                finish(p, s, out);
                continue;
            }
        }
        real_max = std::max(real_max, std::min(hc[C_ITERS] - prev, check));
        M.iters = hc[C_ITERS];
        M.broke = hc[C_BREAK];
        M.done = hc[C_DONE] != 0 || hc[C_STOP] != 0;
        const double rr = hs[S_RR];
        const double ratio = M.o.replace_ratio > 0.0 ? M.o.replace_ratio : 0.1;
        // mixed: replace when the recursive residual fell by `ratio`, and
        // also when it rose far above the last true residual or no
        // replacement happened for a while, so that a refinement that does
        // not contract is seen (and abandoned) early
        const bool mixed_due = M.mixed && (rr <= ratio * ratio * M.rr_ref || rr > kGrow * M.rr_ref ||
                                           M.iters - M.last_rep >= kForceEvery);
        if (!(M.done || mixed_due)) {
            if (M.broke || M.iters >= M.o.maxiter) finish(p, s, out);
            continue;
        }
        M.rr_rec = rr;
        rep.push_back(s);
    }
    if (p->timing) {
        // the SpMVs of iterations in which some member was active executed
        for (int j = 0; j < real_max; j++) {
            float ms = 0.f;
            AMR_HIP(hipEventElapsedTime(&ms, p->ev[2 * j], p->ev[2 * j + 1]));
            p->t_spmv_ms += ms;
            p->n_spmv++;
        }
    }
    if (rep.empty()) return AMR_OK;

    // rt = b - K x in FP64 for the replacing members
    if (int rc = pool_begin_host_writes(p)) return rc;
    build_groups(p, rep, true, ent, gr);
    std::vector<int32_t> sr(rep.begin(), rep.end());
    need = padded(sr);
    for (const auto &e : ent) need += padded(e);
    if (int rc = pool_reserve_lists(p, need)) return rc;
    Ls.off = 0;
    const size_t o_sr = Ls.push(sr);
    {
        int gi = 0;
        for (const auto &e : ent)
            if (!e.empty()) gr[gi++].off = Ls.push(e);
    }
    if (int rc = upload(p, Ls.off, false)) return rc;
    for (const Group &g : gr) launch_residual(g.lanes, g.n, g.Rmax, st, p->dD, Ls.dev_at<int2>(g.off));
    hipLaunchKernelGGL(pcg_finalize_kernel, dim3((unsigned)sr.size()), dim3(WGV), 0, st, p->dD,
                       Ls.dev_at<int32_t>(o_sr), 1);
    AMR_HIP(hipGetLastError());
    AMR_HIP(hipMemcpyAsync(p->hS, p->dS, sbytes, hipMemcpyDeviceToHost, st));
    if (int rc = pool_sync(p)) return rc;

    std::vector<int> cont[2];  // resume as is / with a restarted search direction
    bool desc_changed = false;
    for (int s : rep) {
        Member &M = p->mem[s];
        const double *hs = p->hS + (size_t)s * SEG_DBL;
        M.replacements++;
        M.last_rep = M.iters;
        const double rr = hs[S_RR];
        M.relres = M.bb > 0.0 ? std::sqrt(rr / M.bb) : 0.0;
        const double floor2 = kFloorC * kFloorC * kEps * kEps * hs[S_KX];
        if (rr <= M.o.rtol * M.o.rtol * M.bb || rr <= floor2) {
            M.converged = 1;
            finish(p, s, out);
            continue;
        }
        if (!std::isfinite(rr) || M.iters >= M.o.maxiter) {
            finish(p, s, out);
            continue;
        }
        bool restart = M.broke != 0;
        const double drift = M.o.drift_limit > 0.0 ? M.o.drift_limit : 0.3;
        if (M.mixed && (hs[S_GAP] > drift * drift * std::max(rr, M.rr_rec) || rr > M.rr_true_prev)) {
            // the FP32 operator is too coarse for this model: the recursive
            // residual drifted from the true one, or the refinement stopped
            // contracting; finish on the FP64 operator
            M.mixed = false;
            M.fell_back = 1;
            restart = true;
        } else if ((M.done && !M.broke && rr > 0.25 * M.rr_true_prev) || (M.broke && ++M.breakdowns > 3)) {
            finish(p, s, out);  // stagnation above both bounds, or repeated breakdown
            continue;
        }
        SegDesc &d = p->hD[s];
        const int f32 = M.mixed ? 1 : 0, wp32 = (M.mixed && M.f32x) ? 1 : 0;
        if (d.f32 != f32 || d.wp32 != wp32) {
            d.f32 = f32;
            d.wp32 = wp32;
            desc_changed = true;
        }
        M.broke = 0;
        M.rr_true_prev = rr;
        M.rr_ref = rr;
        cont[restart ? 1 : 0].push_back(s);
    }
    if (cont[0].empty() && cont[1].empty()) return AMR_OK;
    // resume: r = rt, z = M r (p = z on restart), rz, clear the stop flags
    std::vector<int2> vc0 = vec_list(p, cont[0]), vc1 = vec_list(p, cont[1]);
    std::vector<int32_t> sc(cont[0].begin(), cont[0].end());
    sc.insert(sc.end(), cont[1].begin(), cont[1].end());
    std::vector<int> call(sc.begin(), sc.end());
    std::vector<int2> ac = agg_list(p, call), cc = cw_list(p, call), pc = vec_list(p, tl_only(p, cont[1]));
    if (int rc = pool_reserve_lists(p, padded(vc0) + padded(vc1) + padded(sc) + padded(ac) + padded(cc) + padded(pc)))
        return rc;
    Ls.off = 0;
    const size_t o_vc0 = Ls.push(vc0), o_vc1 = Ls.push(vc1), o_sc = Ls.push(sc);
    const size_t o_ac = Ls.push(ac), o_cc = Ls.push(cc), o_pc = Ls.push(pc);
    if (int rc = upload(p, Ls.off, desc_changed || p->desc_dirty)) return rc;
    if (!vc0.empty())
        hipLaunchKernelGGL(pcg_rez_kernel, dim3((unsigned)vc0.size()), dim3(WGV), 0, st, p->dD,
                           Ls.dev_at<int2>(o_vc0), 0);
    if (!vc1.empty())
        hipLaunchKernelGGL(pcg_rez_kernel, dim3((unsigned)vc1.size()), dim3(WGV), 0, st, p->dD,
                           Ls.dev_at<int2>(o_vc1), 1);
    launch_tl_fresh(p, Ls, o_ac, (unsigned)ac.size(), o_cc, (unsigned)cc.size(), o_pc, (unsigned)pc.size());
    hipLaunchKernelGGL(pcg_finalize_kernel, dim3((unsigned)sc.size()), dim3(WGV), 0, st, p->dD,
                       Ls.dev_at<int32_t>(o_sc), 2);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

}  // namespace

int amr_pool_step_collect(amr_pool *p, std::vector<amr_pool_result> &out)
{
    if (!p->in_flight) return AMR_OK;
    AMR_HIP(hipSetDevice(p->device));
    AMR_HIP(hipEventSynchronize(p->ev_step));
    p->in_flight = false;
    const size_t first = out.size();
    const int rc = step_decide(p, out);
    for (size_t i = first; i < out.size(); i++)
        if (out[i].m->stream != out[i].m->own_stream) p->leaving.push_back(out[i].m);
    return rc;
}

// The models that left the pool at its last step return to their own
// streams, after the work queued for them on the pool's stream so far.
int amr_pool_hand_back(amr_pool *p)
{
    if (p->leaving.empty()) return AMR_OK;
    AMR_HIP(hipSetDevice(p->device));
    for (amr_model *m : p->leaving)
        if (m->stream == p->stream && m->pool == p && !amr_pool_is_member(m))
            if (int rc = handoff(m, p->stream, m->own_stream)) return rc;
    p->leaving.clear();
    return AMR_OK;
}

int amr_pool_step_members(const amr_pool *p) { return p->in_flight ? (int)p->fl_act.size() : 0; }

int amr_pool_step_query(amr_pool *p, int *done)
{
    *done = 1;
    if (!p->in_flight) return AMR_OK;
    AMR_HIP(hipSetDevice(p->device));
    const hipError_t q = hipEventQuery(p->ev_step);
    if (q == hipErrorNotReady) {
        (void)hipGetLastError();  // not an error: clear it
        *done = 0;
        return AMR_OK;
    }
    AMR_HIP(q);
    return AMR_OK;
}

int amr_pool_run(amr_pool *p, std::vector<amr_pool_result> &out)
{
    if (int rc = amr_pool_step_launch(p)) return rc;
    return amr_pool_step_collect(p, out);
}

int amr_pool_set_timing(amr_pool *p, int enable, double *ms, int64_t *n)
{
    if (enable && !p->timing) {
        p->t_spmv_ms = 0.0;
        p->n_spmv = 0;
    }
    p->timing = enable != 0;
    if (ms) *ms = p->t_spmv_ms;
    if (n) *n = p->n_spmv;
    return AMR_OK;
}

int amr_pool_graphs(amr_pool *p, int enable, int64_t *captures, int64_t *replays)
{
    if (enable >= 0) {
        if (p->in_flight) {
            amr_set_error("amr_pool_graphs: a step is in flight");
            return AMR_ERR_ARG;
        }
        p->use_graphs = enable != 0;
    }
    if (captures) *captures = p->graph_captures;
    if (replays) *replays = p->graph_replays;
    return AMR_OK;
}

int amr_pool_counts(amr_pool *p, int64_t *out)
{
    out[0] = p->coarse_batches;
    out[1] = p->coarse_formed;
    out[2] = p->iter_kernels;
    out[3] = p->iter_submits;
    out[4] = p->member_kernels;
    return AMR_OK;
}

// ------------------------------------------------- single-model wrappers

namespace {

int self_pool(amr_model *m, amr_pool **out)
{
    if (!m->self_pool) {
        if (int rc = amr_pool_new(&m->self_pool, m->device, 1, 32, m->stream)) return rc;
    }
    m->self_pool->stream = m->stream;
    *out = m->self_pool;
    return AMR_OK;
}

// SpMV lanes per block row.  Measured on MI355X: alone, a model's SpMV is
// fastest with 8 lanes below about 0.75M blocks and with 32 above 4M blocks,
// the choices in between within 10% of each other (64 lanes never best); in
// a pool, members with the same lanes and operand kind share one launch per
// iteration, and one lanes value for all (16) beat the per-model best by
// 8-10% on both pooled workloads (256 Tet10 cantilevers, 48 Hex8/Tet10
// boxes).  So: 16 lanes, 32 for the FP64 operator from 0.75M blocks (16
// lanes cost 20% more there) and for every operator from 4M blocks (the
// 811,923-DOF Hex20 block: 10% less time than with 16).
int choose_lanes(const amr_model *m, int kind)
{
    const int64_t nb = m->pat.nblk;
    return (nb >= 4 * 1024 * 1024 || (kind == K_F64 && nb >= 768 * 1024)) ? 32 : 16;
}

}  // namespace

// y = K x (mask_mode 0) or mask .* (K x) (mask_mode 1), FP64, node order.
int amr_spmv_full(amr_model *m, const double *x, double *y, int mask_mode)
{
    return amr_spmv_vals(m, m->Kv, x, y, mask_mode);
}

int amr_spmv_vals(amr_model *m, const double *vals, const double *x, double *y, int mask_mode)
{
    if (int rc = amr_choose_lanes(m)) return rc;
    const Geom g = model_geom(m);
    const size_t lds = tile_lds(g.R, 0);
#define AMR_SPMV(LN)                                                                                               \
    hipLaunchKernelGGL((spmv_kernel<LN>), dim3((unsigned)g.ntiles), dim3(WGS), lds, m->stream, m->nnodes, g.R,     \
                       m->rowptr, m->col, vals, x, y, m->mask, mask_mode)
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

// Lanes per row for the three operand kinds, unless set explicitly.  Every
// choice gives bit-identical results (see row_sum); the choice depends only
// on the matrix, so it needs no device work.
int amr_choose_lanes(amr_model *m)
{
    if (!m->lanes64) m->lanes64 = choose_lanes(m, K_F64);
    if (!m->lanes32) m->lanes32 = choose_lanes(m, K_F32);
    if (!m->lanes32d) m->lanes32d = choose_lanes(m, K_F32D);
    return AMR_OK;
}

// Time every lanes choice for the three operand kinds on this model's matrix
// and keep the fastest (explicit tuning; amr_choose_lanes is the default).
int amr_tune_lanes(amr_model *m)
{
    if (amr_pool_is_member(m)) {
        amr_set_error("the model is solving in a pool; tune it before adding it");
        return AMR_ERR_ARG;
    }
    amr_pool *p;
    if (int rc = self_pool(m, &p)) return rc;
    if (int rc = pool_begin_host_writes(p)) return rc;
    // a descriptor whose vectors are scratch: p = 1, q, x, r written freely
    Member &M = p->mem[0];
    M = Member();
    M.m = m;
    M.o.maxiter = 1 << 30;
    M.o.rtol = 1.0;
    M.b = m->v[5];
    M.x = m->v[6];
    fill_desc(p, 0);
    M = Member();  // not a solving member
    const Geom g = model_geom(m);
    std::vector<int2> tl;
    for (int t = 0; t < g.ntiles; t++) tl.push_back(make_int2(0, t));
    if (int rc = pool_reserve_lists(p, padded(tl))) return rc;
    Lists Ls{p};
    const size_t o = Ls.push(tl);
    if (int rc = upload(p, Ls.off, true)) return rc;
    p->desc_dirty = true;  // the next real member rewrites it
    const int64_t n = m->nnodes * 3;
    const unsigned gv = (unsigned)g.nvblk;
    AMR_HIP(hipMemsetAsync(p->dS, 0, sizeof(double) * SEG_DBL, p->stream));
    hipLaunchKernelGGL(fill_kernel, dim3(gv), dim3(WGV), 0, p->stream, n, m->v[3], 1.0);
    hipLaunchKernelGGL(fill_f32_kernel, dim3(gv), dim3(WGV), 0, p->stream, n, m->f[0], 1.0f);
    AMR_HIP(hipGetLastError());
    hipEvent_t e0, e1;
    AMR_HIP(hipEventCreate(&e0));
    AMR_HIP(hipEventCreate(&e1));
    int *dst[K_NKIND] = {&m->lanes64, &m->lanes32, &m->lanes32d};
    for (int kind = 0; kind < K_NKIND; kind++) {
        float best = 1e30f;
        int bestL = 64;
        for (int L : LANES) {
            launch_pcg_spmv(kind, L, (unsigned)g.ntiles, g.R, p->stream, p->dD, Ls.dev_at<int2>(o));  // warm
            AMR_HIP(hipEventRecord(e0, p->stream));
            for (int rep = 0; rep < 3; rep++)
                launch_pcg_spmv(kind, L, (unsigned)g.ntiles, g.R, p->stream, p->dD, Ls.dev_at<int2>(o));
            AMR_HIP(hipEventRecord(e1, p->stream));
            AMR_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            AMR_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                bestL = L;
            }
        }
        *dst[kind] = bestL;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    p->pending = false;  // the event waits above drained the stream
    return AMR_OK;
}

// Solve K11 x = b (node order, b zero at prescribed dofs), x FP64: a pool of
// one, stepped until the model leaves it.
int amr_pcg_solve(amr_model *m, const double *b, double *x, const amr_solve_opts *o, amr_solve_info *info)
{
    if (amr_pool_is_member(m)) {
        amr_set_error("the model is solving in a pool; wait for it to leave before solving it alone");
        return AMR_ERR_ARG;
    }
    amr_pool *p;
    if (int rc = amr_choose_lanes(m)) return rc;
    if (int rc = self_pool(m, &p)) return rc;
    p->check = std::max(1, std::min(o->check_every > 0 ? o->check_every : 32, 256));
    p->timing = m->timing;
    p->t_spmv_ms = 0.0;
    p->n_spmv = 0;
    if (int rc = amr_pool_join(p, m, b, x, o)) return rc;
    std::vector<amr_pool_result> res;
    while (res.empty()) {
        if (int rc = amr_pool_run(p, res)) {
            p->mem[0] = Member();
            return rc;
        }
    }
    if (m->timing) {
        m->t_spmv_ms += p->t_spmv_ms;
        m->n_spmv += p->n_spmv;
    }
    const amr_solve_info &r = res[0].info;
    if (info) *info = r;
    if (!r.converged) {
        char buf[256];
        snprintf(buf, sizeof buf, "PCG did not reach rtol=%.3e (relres %.3e after %d iterations, %d replacements)",
                 o->rtol, r.relres, r.iterations, r.refinements);
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

__global__ void scale_kernel(int64_t n, double a, const double *__restrict__ x, double *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * WGV + threadIdx.x; i < n; i += (int64_t)gridDim.x * WGV) out[i] = a * x[i];
}

// flag = 1 if any entry of v is NaN (flag cleared by the caller)
__global__ void nan_flag_kernel(int64_t n, const double *__restrict__ v, int32_t *__restrict__ flag)
{
    bool any = false;
    for (int64_t i = (int64_t)blockIdx.x * WGV + threadIdx.x; i < n; i += (int64_t)gridDim.x * WGV)
        any = any || v[i] != v[i];
    if (__any(any) && threadIdx.x % 64 == 0) atomicOr(flag, 1);
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
    const int off[4] = {P_A, P_B, P_C, P_D};
    for (int i = threadIdx.x; i < nv; i += WGV)
        for (int k = 0; k < 4; k++) v[k] = fmax(v[k], part[off[k] + i]);
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

int read_scalars(amr_model *m, double *hs, int32_t *hc)
{
    AMR_HIP(hipMemcpyAsync(hs, m->scal, sizeof(double) * S_NSCAL, hipMemcpyDeviceToHost, m->stream));
    AMR_HIP(hipMemcpyAsync(hc, m->scal + S_NSCAL, sizeof(int32_t) * C_NCTL, hipMemcpyDeviceToHost, m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    return AMR_OK;
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

int amr_vec_fill_n(amr_model *m, double *x, int64_t n, double val)
{
    const int64_t g = std::max<int64_t>(1, std::min<int64_t>((n + WGV - 1) / WGV, 4 * MAX_GV));
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)g), dim3(WGV), 0, m->stream, n, x, val);
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

int amr_vec_scale(amr_model *m, double a, const double *x, double *out)
{
    hipLaunchKernelGGL(scale_kernel, dim3(vec_grid(m)), dim3(WGV), 0, m->stream, m->nnodes * 3, a, x, out);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

int amr_vec_nan_flag(amr_model *m, const double *v, int32_t *flag)
{
    AMR_HIP(hipMemsetAsync(flag, 0, sizeof(int32_t), m->stream));
    hipLaunchKernelGGL(nan_flag_kernel, dim3(vec_grid(m)), dim3(WGV), 0, m->stream, m->nnodes * 3, v, flag);
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

// The same as amr_vec_nr_post without waiting: the scalars are copied into
// hs/hc (pinned host memory) in stream order
int amr_vec_nr_post_async(amr_model *m, double *hs, int32_t *hc)
{
    double **v = m->sv;
    int32_t *ctl = (int32_t *)(m->scal + S_NSCAL);
    const unsigned g = vec_grid(m);
    hipLaunchKernelGGL(nr_post_kernel, dim3(g), dim3(WGV), 0, m->stream, m->nnodes * 3, m->mask, v[SV_DFEX],
                       v[SV_DFIN], v[SV_DUI], v[SV_DUA], v[SV_R], m->part);
    hipLaunchKernelGGL(nr_finalize_kernel, dim3(1), dim3(WGV), 0, m->stream, (int)g, m->part, ctl, m->scal);
    AMR_HIP(hipGetLastError());
    AMR_HIP(hipMemcpyAsync(hs, m->scal, sizeof(double) * S_NSCAL, hipMemcpyDeviceToHost, m->stream));
    AMR_HIP(hipMemcpyAsync(hc, m->scal + S_NSCAL, sizeof(int32_t) * C_NCTL, hipMemcpyDeviceToHost, m->stream));
    return AMR_OK;
}

int amr_nr_post_slots(int *res, int *errn, int *errd, int *nan)
{
    *res = S_RES;
    *errn = S_ERRN;
    *errd = S_ERRD;
    *nan = C_NAN;
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
