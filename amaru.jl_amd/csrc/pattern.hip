// Device construction of the node-block sparsity structure and of the
// deterministic gather maps (the same arrays as the host builder in
// model_build.cpp, bit for bit):
//
//   fptr/fsrc  node -> (e*nn + a) incidence, ascending element order (the
//              internal-force gather, src/mech/mech-solver.jl:129-136)
//   rowptr/col block rows: the sorted distinct nodes sharing an element with
//              the row's node; diag the block index of (A, A)
//   cptr/csrc  for every block (A, B), the element blocks (e, a, b) that
//              land in it, ascending e (the triplet push!/sparse() sums of
//              src/mech/mech-solver.jl:86-102 in element order)
//
// Row A's contributions are the pairs (i, b) of its incidences i (ascending
// e) and local nodes b: ninc(A) * nn of them, and the rows' lists are
// contiguous in csrc (row A starts at fptr[A] * nn).  One workgroup per row
// sorts the keys (column, position) in LDS: equal columns keep ascending
// position, i.e. ascending element, so the sorted list is the row's csrc
// and its runs of equal column are the blocks.  The first pass counts the
// blocks of each row, a scan gives rowptr, the second pass writes the
// arrays.  No atomics decide an order: the incidence lists are filled with
// atomics and then sorted per node.
#include "amr_internal.h"

#define TRY(x)               \
    do {                     \
        int _rc = (x);       \
        if (_rc) return _rc; \
    } while (0)

namespace {

constexpr int SCAN_T = 256, SCAN_V = 8, SCAN_B = SCAN_T * SCAN_V;  // items per scan block
constexpr int ROW_T = 256;     // threads of a row workgroup
constexpr int ROW_CAP = 2048;  // contributions of a row sorted in LDS

// Exclusive prefix sums of SCAN_B-item blocks; bsum[b] = block b's total.
// Each thread reads its items before writing them, so in == out is allowed.
__global__ void __launch_bounds__(SCAN_T) scan_block_kernel(int64_t n, const int32_t *in, int32_t *out,
                                                            int32_t *__restrict__ bsum)
{
    __shared__ int32_t s[SCAN_T];
    const int64_t base = (int64_t)blockIdx.x * SCAN_B + (int64_t)threadIdx.x * SCAN_V;
    int32_t v[SCAN_V], t = 0;
#pragma unroll
    for (int i = 0; i < SCAN_V; i++) {
        v[i] = base + i < n ? in[base + i] : 0;
        t += v[i];
    }
    s[threadIdx.x] = t;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        const int32_t a = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += a;
        __syncthreads();
    }
    int32_t run = s[threadIdx.x] - t;
#pragma unroll
    for (int i = 0; i < SCAN_V; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == SCAN_T - 1) bsum[blockIdx.x] = s[SCAN_T - 1];
}

__global__ void __launch_bounds__(256) scan_add_kernel(int64_t n, int32_t *__restrict__ out,
                                                       const int32_t *__restrict__ boff)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += boff[i / SCAN_B];
}

// In-place exclusive scan of x[0..n); tmp holds at least scan_tmp_ints(n).
int64_t scan_tmp_ints(int64_t n)
{
    int64_t t = 0;
    while (n > SCAN_B) {
        n = (n + SCAN_B - 1) / SCAN_B;
        t += n;
    }
    return t + 1;
}

void exclusive_scan(int32_t *x, int64_t n, int32_t *tmp, hipStream_t s)
{
    if (n <= 0) return;
    const int64_t nb = (n + SCAN_B - 1) / SCAN_B;
    hipLaunchKernelGGL(scan_block_kernel, dim3((unsigned)nb), dim3(SCAN_T), 0, s, n, x, x, tmp);
    if (nb == 1) return;
    exclusive_scan(tmp, nb, tmp + nb, s);
    hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, x, tmp);
}

// flags[0]: a connectivity entry outside [0, nnodes) (flags[1] its value);
// flags[2]: a row with more than ROW_CAP contributions; flags[3]: a block
// whose column offset does not fit 16 bits
__global__ void __launch_bounds__(256) inc_count_kernel(int64_t nent, int64_t nnodes, const int32_t *__restrict__ conn,
                                                        int32_t *__restrict__ cnt, int32_t *__restrict__ flags)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nent) return;
    const int32_t n = conn[i];
    if (n < 0 || n >= nnodes) {
        atomicMax(&flags[0], 1);
        atomicExch(&flags[1], n);
        return;
    }
    atomicAdd(&cnt[n], 1);
}

__global__ void __launch_bounds__(256) inc_fill_kernel(int64_t nent, int64_t nnodes, const int32_t *__restrict__ conn,
                                                       const int32_t *__restrict__ ninc, int32_t *__restrict__ fill,
                                                       int32_t *__restrict__ inc)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nent) return;
    const int32_t n = conn[i];
    if (n < 0 || n >= nnodes) return;
    const int32_t p = atomicAdd(&fill[n], 1);
    inc[ninc[n] + p] = (int32_t)i;   // e*nn + a
}

// ascending (e*nn + a) within each node's incidence list
__global__ void __launch_bounds__(256) inc_sort_kernel(int64_t nnodes, const int32_t *__restrict__ ninc,
                                                       int32_t *__restrict__ inc)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nnodes) return;
    const int32_t b = ninc[n], e = ninc[n + 1];
    for (int32_t i = b + 1; i < e; i++) {
        const int32_t v = inc[i];
        int32_t j = i - 1;
        while (j >= b && inc[j] > v) {
            inc[j + 1] = inc[j];
            j--;
        }
        inc[j + 1] = v;
    }
}

// PASS 0: rowlen[A] = blocks of row A.  PASS 1: csrc, col, cptr, diag and
// col16 of row A (rowptr scanned).
template <int PASS>
__global__ void __launch_bounds__(ROW_T) row_kernel(int64_t nnodes, int nn, const int32_t *__restrict__ conn,
                                                    const int32_t *__restrict__ ninc, const int32_t *__restrict__ inc,
                                                    int32_t *__restrict__ rowptr, int32_t *__restrict__ csrc,
                                                    int32_t *__restrict__ col, int64_t *__restrict__ cptr,
                                                    int32_t *__restrict__ diag, int16_t *__restrict__ col16,
                                                    int32_t *__restrict__ flags)
{
    __shared__ uint64_t key[ROW_CAP];
    __shared__ int32_t part[ROW_T];
    const int64_t A = blockIdx.x;
    const int tid = threadIdx.x;
    const int32_t i0 = ninc[A];
    const int M = (ninc[A + 1] - i0) * nn;
    if (M > ROW_CAP) {
        if (tid == 0) {
            atomicMax(&flags[2], 1);
            if (PASS == 0) rowptr[A] = 0;
        }
        return;
    }
    int P = 1;
    while (P < M) P <<= 1;
    for (int t = tid; t < P; t += ROW_T) {
        uint64_t k = ~0ull;
        if (t < M) {
            const int32_t ea = inc[i0 + t / nn];
            const int32_t B = conn[(int64_t)(ea / nn) * nn + t % nn];
            k = ((uint64_t)(uint32_t)B << 32) | (uint32_t)t;
        }
        key[t] = k;
    }
    __syncthreads();
    // bitonic sort, ascending
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = tid; t < P; t += ROW_T) {
                const int u = t ^ j;
                if (u > t) {
                    const uint64_t a = key[t], b = key[u];
                    if (((t & k) == 0) == (a > b)) {
                        key[t] = b;
                        key[u] = a;
                    }
                }
            }
            __syncthreads();
        }
    // block starts: each thread scans a contiguous run of positions
    const int C = (M + ROW_T - 1) / ROW_T;
    const int lo = tid * C < M ? tid * C : M, hi = lo + C < M ? lo + C : M;
    int nst = 0;
    for (int t = lo; t < hi; t++) nst += (t == 0 || (key[t] >> 32) != (key[t - 1] >> 32));
    part[tid] = nst;
    __syncthreads();
    for (int off = 1; off < ROW_T; off <<= 1) {
        const int32_t a = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += a;
        __syncthreads();
    }
    if (PASS == 0) {
        if (tid == ROW_T - 1) rowptr[A] = part[ROW_T - 1];
        return;
    }
    const int64_t base = (int64_t)i0 * nn;
    const int32_t r0 = rowptr[A];
    int32_t j = r0 + part[tid] - nst;
    for (int t = lo; t < hi; t++) {
        const uint64_t k = key[t];
        const int32_t B = (int32_t)(k >> 32), pos = (int32_t)(uint32_t)k;
        csrc[base + t] = inc[i0 + pos / nn] * nn + pos % nn;   // (e*nn + a)*nn + b
        if (t == 0 || (key[t - 1] >> 32) != (k >> 32)) {
            col[j] = B;
            cptr[j] = base + t;
            if (B == A) diag[A] = j;
            const int64_t d = (int64_t)B - A;
            if (d < -32768 || d > 32767) atomicMax(&flags[3], 1);
            else if (col16) col16[j] = (int16_t)d;
            j++;
        }
    }
}

__global__ void fill_kernel(int64_t n, int32_t *__restrict__ x, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

inline unsigned blocks(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

// Builds the pattern of a model whose conn is on the device.  Uses the
// model's fptr (incidence counts, then offsets), fsrc (incidences), rowptr
// (block counts, then offsets), diag (fill counters, then diagonal) and
// part (scan partials and flags) as they are allocated; allocates col, cptr,
// csrc and col16.  Returns AMR_PATTERN_ON_HOST (having allocated nothing)
// when a row has more contributions than one workgroup sorts: the host
// builder serves those models.
int amr_build_pattern_device(amr_model *m)
{
    const int64_t n = m->nnodes, nn = m->nn, nent = m->ne * nn;
    hipStream_t s = m->stream;
    if (m->ne * nn * nn > INT32_MAX) {
        amr_set_error("model too large for 32-bit element-block ids");
        return AMR_ERR_ARG;
    }
    int32_t *flags = (int32_t *)m->part, *stmp = flags + 4;
    if ((int64_t)(AMR_PART_DBL * 2 - 4) < scan_tmp_ints(n + 1)) {
        amr_set_error("model too large for the pattern scan");
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipMemsetAsync(flags, 0, 4 * sizeof(int32_t), s));
    AMR_HIP(hipMemsetAsync(m->fptr, 0, (n + 1) * sizeof(int32_t), s));
    AMR_HIP(hipMemsetAsync(m->diag, 0, n * sizeof(int32_t), s));
    if (nent > 0) hipLaunchKernelGGL(inc_count_kernel, dim3(blocks(nent)), dim3(256), 0, s, nent, n, m->conn, m->fptr, flags);
    exclusive_scan(m->fptr, n + 1, stmp, s);
    if (nent > 0) {
        hipLaunchKernelGGL(inc_fill_kernel, dim3(blocks(nent)), dim3(256), 0, s, nent, n, m->conn, m->fptr, m->diag,
                           m->fsrc);
        hipLaunchKernelGGL(inc_sort_kernel, dim3(blocks(n)), dim3(256), 0, s, n, m->fptr, m->fsrc);
    }
    AMR_HIP(hipMemsetAsync(m->rowptr + n, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(row_kernel<0>, dim3((unsigned)n), dim3(ROW_T), 0, s, n, (int)nn, m->conn, m->fptr, m->fsrc,
                       m->rowptr, nullptr, nullptr, nullptr, nullptr, nullptr, flags);
    AMR_HIP(hipGetLastError());
    exclusive_scan(m->rowptr, n + 1, stmp, s);
    int32_t h[4], nblk = 0;
    AMR_HIP(hipMemcpyAsync(h, flags, sizeof(h), hipMemcpyDeviceToHost, s));
    AMR_HIP(hipMemcpyAsync(&nblk, m->rowptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AMR_HIP(hipStreamSynchronize(s));
    if (h[0]) {
        amr_set_error("connectivity references node " + std::to_string(h[1]) + " outside [0, " + std::to_string(n) +
                      ")");
        return AMR_ERR_ARG;
    }
    if (h[2]) return AMR_PATTERN_ON_HOST;
    m->pat.nnodes = n;
    m->pat.nblk = nblk;
    m->pat.ncontrib = nent * nn;
    TRY(amr_dalloc(m, (void **)&m->col, sizeof(int32_t) * (size_t)nblk));
    TRY(amr_dalloc(m, (void **)&m->cptr, sizeof(int64_t) * (size_t)(nblk + 1)));
    TRY(amr_dalloc(m, (void **)&m->csrc, sizeof(int32_t) * (size_t)m->pat.ncontrib));
    TRY(amr_dalloc(m, (void **)&m->col16, sizeof(int16_t) * (size_t)nblk));
    hipLaunchKernelGGL(fill_kernel, dim3(blocks(n)), dim3(256), 0, s, n, m->diag, -1);
    hipLaunchKernelGGL(row_kernel<1>, dim3((unsigned)n), dim3(ROW_T), 0, s, n, (int)nn, m->conn, m->fptr, m->fsrc,
                       m->rowptr, m->csrc, m->col, m->cptr, m->diag, m->col16, flags);
    AMR_HIP(hipGetLastError());
    const int64_t total = m->pat.ncontrib;
    AMR_HIP(hipMemcpyAsync(m->cptr + nblk, &total, sizeof(int64_t), hipMemcpyHostToDevice, s));
    AMR_HIP(hipMemcpyAsync(h, flags, sizeof(h), hipMemcpyDeviceToHost, s));
    AMR_HIP(hipStreamSynchronize(s));
    if (h[3]) {
        amr_release(m, m->col16);
        m->col16 = nullptr;
    }
    return AMR_OK;
}
