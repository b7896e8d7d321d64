// Structured block meshes and DOF maps on the device: split_block! for
// HEX8, HEX20 and TET10 (src/mesh/structured.jl:384-459, 461-552) with the
// synchronize! numbering (src/mesh/mesh.jl:348-350), and the DOF numbering of
// configure_dofs! (src/bc.jl:198-233).  Same arithmetic as the host mesher,
// so coordinates, connectivity and eq ids are bit-identical to it.
//
// Grid points of a block: (ni, nj, nk) = (sub*n + 1) per axis, sub = 1 for
// HEX8 and 2 otherwise, k outer, j, i inner.  HEX20 keeps the points with at
// most one odd (0-based) index; node ids count kept points in creation order
// and have a closed form per (i, j, k).  Natural coordinates
// r = -1 + 2 ((1 / (sub nx)) i), physical ones the HEX8 map of the corners
// summed corner by corner, rounded to 8 decimals (round half to even of
// x * 1e8, -0 dropped: src/node.jl:58-60).
#include <algorithm>
#include <mutex>
#include <vector>

#include "amr_internal.h"

namespace {

__constant__ double kSigns[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                    {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
// grid offsets of the 27 points of a 2x2x2 sub-grid cell (points p1..p27 of
// src/mesh/structured.jl:499-534), and the six tetrahedra of a cell as
// 0-based indices of those points (src/mesh/structured.jl:537-542)
__constant__ int kP27[27][3] = {{0, 0, 0}, {2, 0, 0}, {2, 2, 0}, {0, 2, 0}, {0, 0, 2}, {2, 0, 2}, {2, 2, 2}, {0, 2, 2}, {1, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 1, 0}, {1, 0, 2}, {2, 1, 2}, {1, 2, 2}, {0, 1, 2}, {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1}, {0, 1, 1}, {2, 1, 1}, {1, 0, 1}, {1, 2, 1}, {1, 1, 0}, {1, 1, 2}, {1, 1, 1}};
__constant__ int kTetSplit[6][10] = {
    {1, 3, 0, 7, 24, 11, 8, 26, 19, 20},
    {1, 0, 4, 7, 8, 16, 22, 26, 20, 15},
    {1, 4, 5, 7, 22, 12, 17, 26, 15, 25},
    {1, 5, 6, 7, 17, 13, 21, 26, 25, 14},
    {1, 2, 3, 7, 9, 10, 24, 26, 23, 19},
    {1, 6, 2, 7, 21, 18, 9, 26, 14, 23}};
__constant__ int kHex8Off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                   {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

struct Grid {
    int64_t ni, nj, nk;
    int sub;
    bool skip;   // HEX20: drop points with two or more odd indices
};

__host__ __device__ inline bool kept(const Grid &g, int64_t i, int64_t j, int64_t k)
{
    if (!g.skip) return true;
    const int odd = (int)(i & 1) + (int)(j & 1) + (int)(k & 1);
    return odd <= 1;
}

// kept points in row (j, k), and in layer k, of a HEX20 grid
__host__ __device__ inline int64_t row_count(const Grid &g, int64_t j, int64_t k)
{
    const int odd = (int)(j & 1) + (int)(k & 1);
    return odd == 0 ? g.ni : odd == 1 ? (g.ni + 1) / 2 : 0;
}

__host__ __device__ inline int64_t layer_count(const Grid &g, int64_t k)
{
    const int64_t je = (g.nj + 1) / 2, jo = g.nj / 2;   // even and odd rows
    return (k & 1) ? je * ((g.ni + 1) / 2) : je * g.ni + jo * ((g.ni + 1) / 2);
}

// node id of a kept grid point
__host__ __device__ inline int64_t node_id(const Grid &g, int64_t i, int64_t j, int64_t k)
{
    if (!g.skip) return (k * g.nj + j) * g.ni + i;
    const int64_t ke = (k + 1) / 2, ko = k / 2;   // even and odd layers below k
    int64_t id = ke * layer_count(g, 0) + ko * layer_count(g, 1);
    const int64_t je = (j + 1) / 2, jo = j / 2;   // even and odd rows below j in layer k
    id += je * row_count(g, 0, k) + jo * row_count(g, 1, k);
    id += row_count(g, j, k) == g.ni ? i : (i + 1) / 2;
    return id;
}

__host__ __device__ inline int64_t grid_nodes(const Grid &g)
{
    if (!g.skip) return g.ni * g.nj * g.nk;
    const int64_t ke = (g.nk + 1) / 2, ko = g.nk / 2;
    return ke * layer_count(g, 0) + ko * layer_count(g, 1);
}

__global__ void __launch_bounds__(256) grid_points_kernel(Grid g, int64_t nx, int64_t ny, int64_t nz,
                                                          const double *__restrict__ corners,
                                                          double *__restrict__ X)
{
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= g.ni * g.nj * g.nk) return;
    const int64_t i = p % g.ni, j = (p / g.ni) % g.nj, k = p / (g.ni * g.nj);
    if (!kept(g, i, j, k)) return;
    const double r = -1.0 + 2.0 * ((1.0 / (double)(g.sub * nx)) * (double)i);
    const double s = -1.0 + 2.0 * ((1.0 / (double)(g.sub * ny)) * (double)j);
    const double t = -1.0 + 2.0 * ((1.0 / (double)(g.sub * nz)) * (double)k);
    double x[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 8; a++) {
        const double na = 0.125 * (1.0 + kSigns[a][0] * r) * (1.0 + kSigns[a][1] * s) * (1.0 + kSigns[a][2] * t);
        for (int c = 0; c < 3; c++) x[c] += na * corners[a * 3 + c];
    }
    const int64_t id = node_id(g, i, j, k);
    for (int c = 0; c < 3; c++) X[id * 3 + c] = rint(x[c] * 1e8) / 1e8 + 0.0;
}

__global__ void __launch_bounds__(256) cells_kernel(Grid g, int etype, int64_t nx, int64_t ny, int64_t nz,
                                                    int32_t *__restrict__ conn)
{
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= nx * ny * nz) return;
    const int64_t ci = c % nx, cj = (c / nx) % ny, ck = c / (nx * ny);
    if (etype == AMR_HEX8) {
        for (int a = 0; a < 8; a++)
            conn[c * 8 + a] = (int32_t)node_id(g, ci + kHex8Off[a][0], cj + kHex8Off[a][1], ck + kHex8Off[a][2]);
        return;
    }
    const int64_t i0 = 2 * ci, j0 = 2 * cj, k0 = 2 * ck;
    if (etype == AMR_HEX20) {
        for (int a = 0; a < 20; a++)
            conn[c * 20 + a] = (int32_t)node_id(g, i0 + kP27[a][0], j0 + kP27[a][1], k0 + kP27[a][2]);
        return;
    }
    for (int t = 0; t < 6; t++)
        for (int a = 0; a < 10; a++) {
            const int q = kTetSplit[t][a];
            conn[(c * 6 + t) * 10 + a] = (int32_t)node_id(g, i0 + kP27[q][0], j0 + kP27[q][1], k0 + kP27[q][2]);
        }
}

// DOF numbering: per 256-DOF tile, unknown and prescribed counts, then
// ranks within the tile plus the tile offsets (unknowns first, both in
// DOF order).
__global__ void __launch_bounds__(256) dof_count_kernel(int64_t n3, const uint8_t *__restrict__ active,
                                                        const uint8_t *__restrict__ presc, int32_t *__restrict__ cnt)
{
    __shared__ int su[256], sp[256];
    const int64_t d = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool on = d < n3 && (!active || active[d / 3]);
    su[threadIdx.x] = on && !presc[d < n3 ? d : 0];
    sp[threadIdx.x] = on && presc[d < n3 ? d : 0];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            su[threadIdx.x] += su[threadIdx.x + o];
            sp[threadIdx.x] += sp[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cnt[blockIdx.x * 2] = su[0];
        cnt[blockIdx.x * 2 + 1] = sp[0];
    }
}

// Exclusive tile offsets (unknowns and prescribed, as int64) from the
// per-tile counts, and the totals sizes[0] = ndofs, sizes[1] = nu: one
// workgroup; thread i sums a contiguous run of tiles, a scan over the
// threads gives each run's base.
__global__ void __launch_bounds__(1024) dof_scan_kernel(int64_t nt, const int32_t *__restrict__ cnt,
                                                        int64_t *__restrict__ off, int64_t *__restrict__ sizes)
{
    __shared__ int64_t au[1024], ap[1024];
    const int tid = threadIdx.x;
    const int64_t per = (nt + 1023) / 1024;
    const int64_t t0 = std::min<int64_t>(nt, tid * per), t1 = std::min<int64_t>(nt, t0 + per);
    int64_t su = 0, sp = 0;
    for (int64_t t = t0; t < t1; t++) {
        su += cnt[t * 2];
        sp += cnt[t * 2 + 1];
    }
    au[tid] = su;
    ap[tid] = sp;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int64_t a = tid >= o ? au[tid - o] : 0;
        const int64_t b = tid >= o ? ap[tid - o] : 0;
        __syncthreads();
        au[tid] += a;
        ap[tid] += b;
        __syncthreads();
    }
    int64_t bu = au[tid] - su, bp = ap[tid] - sp;
    for (int64_t t = t0; t < t1; t++) {
        off[t * 2] = bu;
        off[t * 2 + 1] = bp;
        bu += cnt[t * 2];
        bp += cnt[t * 2 + 1];
    }
    if (tid == 1023) {
        sizes[0] = au[1023] + ap[1023];
        sizes[1] = au[1023];
    }
}

__global__ void __launch_bounds__(256) dof_rank_kernel(int64_t n3, const int64_t *__restrict__ sizes,
                                                       const uint8_t *__restrict__ active,
                                                       const uint8_t *__restrict__ presc,
                                                       const int64_t *__restrict__ off, int32_t *__restrict__ eq)
{
    __shared__ int su[256], sp[256];
    const int64_t d = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool on = d < n3 && (!active || active[d / 3]);
    const bool pr = on && presc[d < n3 ? d : 0];
    su[threadIdx.x] = on && !pr;
    sp[threadIdx.x] = pr;
    __syncthreads();
    // inclusive scans (Hillis-Steele)
    for (int o = 1; o < 256; o <<= 1) {
        const int a = (int)threadIdx.x >= o ? su[threadIdx.x - o] : 0;
        const int b = (int)threadIdx.x >= o ? sp[threadIdx.x - o] : 0;
        __syncthreads();
        su[threadIdx.x] += a;
        sp[threadIdx.x] += b;
        __syncthreads();
    }
    if (d >= n3) return;
    if (!on) {
        eq[d] = -1;
        return;
    }
    const int64_t nu = sizes[1];
    eq[d] = pr ? (int32_t)(nu + off[blockIdx.x * 2 + 1] + sp[threadIdx.x] - 1)
               : (int32_t)(off[blockIdx.x * 2] + su[threadIdx.x] - 1);
}

// The three DOF-map kernels on stream s: cnt (nt*2 int32), off (nt*2
// int64) and sizes (2 int64) are work and result space in device memory.
int launch_dof_map(int64_t n3, const uint8_t *active, const uint8_t *presc, int32_t *eq, int32_t *cnt, int64_t *off,
                   int64_t *sizes, hipStream_t s)
{
    const int64_t nt = (n3 + 255) / 256;
    hipLaunchKernelGGL(dof_count_kernel, dim3((unsigned)nt), dim3(256), 0, s, n3, active, presc, cnt);
    hipLaunchKernelGGL(dof_scan_kernel, dim3(1), dim3(1024), 0, s, nt, cnt, off, sizes);
    hipLaunchKernelGGL(dof_rank_kernel, dim3((unsigned)nt), dim3(256), 0, s, n3, sizes, active, presc, off, eq);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

template <typename T>
int dmalloc(T **p, size_t n)
{
    hipError_t e = hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        amr_set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? AMR_ERR_NOMEM : AMR_ERR_HIP;
    }
    return AMR_OK;
}

struct DevBufs {
    std::vector<void *> p;
    ~DevBufs()
    {
        for (void *q : p)
            if (q) (void)hipFree(q);
    }
};

}  // namespace

extern "C" int amr_block_mesh(int device, int etype, const double *corners, int64_t nx, int64_t ny, int64_t nz,
                              int64_t *nnodes, int64_t *ne, double *coords, int32_t *conn)
{
    if (!corners || !nnodes || !ne || nx < 1 || ny < 1 || nz < 1 ||
        (etype != AMR_HEX8 && etype != AMR_HEX20 && etype != AMR_TET10)) {
        amr_set_error("amr_block_mesh: bad element type, division or null pointer");
        return AMR_ERR_ARG;
    }
    Grid g;
    g.sub = etype == AMR_HEX8 ? 1 : 2;
    g.ni = g.sub * nx + 1;
    g.nj = g.sub * ny + 1;
    g.nk = g.sub * nz + 1;
    g.skip = etype == AMR_HEX20;
    const int64_t nn = grid_nodes(g);
    const int64_t nc = nx * ny * nz * (etype == AMR_TET10 ? 6 : 1);
    const int nen = etype == AMR_HEX8 ? 8 : etype == AMR_HEX20 ? 20 : 10;
    if (nn > INT32_MAX || nc * nen > INT32_MAX) {
        amr_set_error("amr_block_mesh: the mesh exceeds 32-bit node ids");
        return AMR_ERR_ARG;
    }
    *nnodes = nn;
    *ne = nc;
    if (!coords && !conn) return AMR_OK;
    if (!coords || !conn) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(device));
    DevBufs bufs;
    double *dX = nullptr, *dC = nullptr;
    int32_t *dconn = nullptr;
    int rc = dmalloc(&dX, (size_t)nn * 3);
    bufs.p.push_back(dX);
    if (!rc) rc = dmalloc(&dC, 24);
    bufs.p.push_back(dC);
    if (!rc) rc = dmalloc(&dconn, (size_t)nc * nen);
    bufs.p.push_back(dconn);
    if (rc) return rc;
    hipStream_t s = nullptr;
    AMR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    AMR_HIP(hipMemcpyAsync(dC, corners, 24 * sizeof(double), hipMemcpyHostToDevice, s));
    const int64_t np = g.ni * g.nj * g.nk;
    hipLaunchKernelGGL(grid_points_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, g, nx, ny, nz, dC,
                       dX);
    AMR_HIP(hipGetLastError());
    const int64_t ncell = nx * ny * nz;
    hipLaunchKernelGGL(cells_kernel, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s, g, etype, nx, ny, nz,
                       dconn);
    AMR_HIP(hipGetLastError());
    AMR_HIP(hipMemcpyAsync(coords, dX, sizeof(double) * nn * 3, hipMemcpyDeviceToHost, s));
    AMR_HIP(hipMemcpyAsync(conn, dconn, sizeof(int32_t) * nc * nen, hipMemcpyDeviceToHost, s));
    AMR_HIP(hipStreamSynchronize(s));
    return AMR_OK;
}

extern "C" int amr_dof_map(int device, int64_t nnodes, const uint8_t *active, const uint8_t *presc, int32_t *eq,
                           int64_t *ndofs, int64_t *nu)
{
    if (nnodes < 0 || !presc || !eq || !ndofs || !nu) return AMR_ERR_ARG;
    const int64_t n3 = nnodes * 3;
    if (n3 > INT32_MAX) {
        amr_set_error("amr_dof_map: over 2^31 dofs");
        return AMR_ERR_ARG;
    }
    *ndofs = 0;
    *nu = 0;
    if (n3 == 0) return AMR_OK;
    AMR_HIP(hipSetDevice(device));
    const int64_t nt = (n3 + 255) / 256;
    DevBufs bufs;
    uint8_t *dact = nullptr, *dpre = nullptr;
    int32_t *dcnt = nullptr, *deq = nullptr;
    int64_t *doff = nullptr;
    int rc = AMR_OK;
    if (active) {
        rc = dmalloc(&dact, (size_t)nnodes);
        bufs.p.push_back(dact);
    }
    if (!rc) rc = dmalloc(&dpre, (size_t)n3);
    bufs.p.push_back(dpre);
    if (!rc) rc = dmalloc(&dcnt, (size_t)nt * 2);
    bufs.p.push_back(dcnt);
    if (!rc) rc = dmalloc(&doff, (size_t)nt * 2 + 2);   // offsets, then ndofs and nu
    bufs.p.push_back(doff);
    if (!rc) rc = dmalloc(&deq, (size_t)n3);
    bufs.p.push_back(deq);
    if (rc) return rc;
    hipStream_t s = nullptr;
    AMR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    if (active) AMR_HIP(hipMemcpyAsync(dact, active, (size_t)nnodes, hipMemcpyHostToDevice, s));
    AMR_HIP(hipMemcpyAsync(dpre, presc, (size_t)n3, hipMemcpyHostToDevice, s));
    if (int e = launch_dof_map(n3, dact, dpre, deq, dcnt, doff, doff + nt * 2, s)) return e;
    int64_t sizes[2];
    AMR_HIP(hipMemcpyAsync(eq, deq, sizeof(int32_t) * n3, hipMemcpyDeviceToHost, s));
    AMR_HIP(hipMemcpyAsync(sizes, doff + nt * 2, sizeof(sizes), hipMemcpyDeviceToHost, s));
    AMR_HIP(hipStreamSynchronize(s));
    *ndofs = sizes[0];
    *nu = sizes[1];
    return AMR_OK;
}

// ---------------------------------------------------------------------------
// amr_dof_map_noalloc: the same kernels with their work space in a static
// device array instead of allocated buffers.  Layout: per-tile counts
// (AMR_DOFMAP_NOALLOC_MAX_DOFS/256 x 2 int32), then per-tile offsets (same
// count x 2 int64), then ndofs and nu are written to the caller's sizes.
// Every device holds its own copy at its own address, looked up once per
// device with hipGetSymbolAddress; host code uses only that address.
static constexpr int64_t kScrTiles = AMR_DOFMAP_NOALLOC_MAX_DOFS / 256;
static constexpr size_t kScrCnt = (size_t)kScrTiles * 2 * sizeof(int32_t);
static constexpr size_t kScrBytes = kScrCnt + (size_t)kScrTiles * 2 * sizeof(int64_t);
static __device__ __align__(256) unsigned char scratch[kScrBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// the device's address of scratch; call with the device current
int scratch_address(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        AMR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = p;
    }
    *out = (unsigned char *)g_scratch_ptr[device];
    return AMR_OK;
}
}  // namespace

extern "C" int amr_dof_map_noalloc(int device, int64_t nnodes, const uint8_t *active, const uint8_t *presc,
                                   int32_t *eq, int64_t *sizes, void *stream)
{
    if (nnodes < 0 || !presc || !eq || !sizes) {
        amr_set_error("amr_dof_map_noalloc: null pointer or negative node count");
        return AMR_ERR_ARG;
    }
    const int64_t n3 = nnodes * 3;
    if (n3 > AMR_DOFMAP_NOALLOC_MAX_DOFS) {
        amr_set_error("amr_dof_map_noalloc: " + std::to_string(n3) + " dofs exceed the scratch limit of " +
                      std::to_string((int64_t)AMR_DOFMAP_NOALLOC_MAX_DOFS));
        return AMR_ERR_ARG;
    }
    int ndev = 0;
    AMR_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev || device >= kMaxDevices) {
        amr_set_error("amr_dof_map_noalloc: no device " + std::to_string(device));
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipSetDevice(device));
    unsigned char *scr = nullptr;
    if (int rc = scratch_address(device, &scr)) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (n3 == 0) {
        // ndofs = nu = 0 through the scan kernel over no tiles
        hipLaunchKernelGGL(dof_scan_kernel, dim3(1), dim3(1024), 0, s, (int64_t)0, (const int32_t *)scr,
                           (int64_t *)(scr + kScrCnt), sizes);
        AMR_HIP(hipGetLastError());
        return AMR_OK;
    }
    return launch_dof_map(n3, active, presc, eq, (int32_t *)scr, (int64_t *)(scr + kScrCnt), sizes, s);
}
