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

__global__ void __launch_bounds__(256) dof_rank_kernel(int64_t n3, int64_t nu, const uint8_t *__restrict__ active,
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
    eq[d] = pr ? (int32_t)(nu + off[blockIdx.x * 2 + 1] + sp[threadIdx.x] - 1)
               : (int32_t)(off[blockIdx.x * 2] + su[threadIdx.x] - 1);
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
    if (!rc) rc = dmalloc(&doff, (size_t)nt * 2);
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
    hipLaunchKernelGGL(dof_count_kernel, dim3((unsigned)nt), dim3(256), 0, s, n3, dact, dpre, dcnt);
    AMR_HIP(hipGetLastError());
    std::vector<int32_t> cnt((size_t)nt * 2);
    AMR_HIP(hipMemcpyAsync(cnt.data(), dcnt, sizeof(int32_t) * nt * 2, hipMemcpyDeviceToHost, s));
    AMR_HIP(hipStreamSynchronize(s));
    // exclusive tile offsets (host: one value per 256 dofs)
    std::vector<int64_t> off((size_t)nt * 2);
    int64_t au = 0, ap = 0;
    for (int64_t t = 0; t < nt; t++) {
        off[t * 2] = au;
        off[t * 2 + 1] = ap;
        au += cnt[t * 2];
        ap += cnt[t * 2 + 1];
    }
    AMR_HIP(hipMemcpyAsync(doff, off.data(), sizeof(int64_t) * nt * 2, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(dof_rank_kernel, dim3((unsigned)nt), dim3(256), 0, s, n3, au, dact, dpre, doff, deq);
    AMR_HIP(hipGetLastError());
    AMR_HIP(hipMemcpyAsync(eq, deq, sizeof(int32_t) * n3, hipMemcpyDeviceToHost, s));
    AMR_HIP(hipStreamSynchronize(s));
    *ndofs = au + ap;
    *nu = au;
    return AMR_OK;
}
