// Output post-processing on the device: the per-IP output values of the
// solid materials (ip_state_vals over stress_strain_dict,
// src/tools/tensors.jl:162-215; src/mech/mat/linear-elastic.jl:140-142,
// von-mises.jl:159-168, drucker-prager.jl:152-163) and nodal patch recovery
// (nodal_patch_recovery, src/fe-model.jl:506-692).
//
// Patch recovery fits, for every patch (the elements around a corner node),
// the polynomial 1, x, y, z, xy, yz, xz (fewer terms for small patches) to
// the IP values by least squares and evaluates it at the patch's nodes; a
// node's value is the mean over the patches that reach it.  The patches,
// their node lists and the IP coordinates depend on the mesh only and are
// built once per model on the host.  Per call:
//   ipvals_kernel   one thread per IP: 19 output values (AoS, IP-major)
//   patch_fit_kernel one wave per patch: Householder QR of the patch's basis
//                   matrix with the IP values appended (rows in registers,
//                   column products reduced across the wave), back
//                   substitution; the basis is centred on the patch's node and
//                   scaled by the patch size (the same polynomial space,
//                   better conditioned)
//   node_eval_kernel one thread per (node, field): the patches' polynomials at
//                   the node, summed in patch order, divided by their count.
// A model that mixes material kinds has fields that only some elements
// carry (the plastic strain of von Mises, the Drucker-Prager extras): as the
// reference (fe-model.jl:614-680), each such field is fitted on the
// subpatch of the elements that carry it and averaged over the subpatches
// that reach a node.  The subpatches are fit units of their own, after the
// patches; the nodes keep two patch lists, one for the fields every element
// carries and one for the extras.
// Every sum has a fixed order: the results do not depend on the launch.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "amr_internal.h"

namespace {

constexpr int NF = AMR_NFIELDS;
constexpr int NT = 7;             // polynomial terms of a full patch
constexpr int NCOL = NT + NF;     // basis columns, then the IP values
constexpr int MAXM = 256;         // IP rows of one patch
constexpr double SR2 = 1.4142135623730951;

// Facet node lists (0-based) of the element shapes, corners first
// (src/shape/solids3d.jl facet_idxs of HEX8, HEX20, TET10).
const int kHexFacets[6][8] = {{0, 4, 7, 3, 16, 15, 19, 11}, {1, 2, 6, 5, 9, 18, 13, 17},
                              {0, 1, 5, 4, 8, 17, 12, 16},  {2, 3, 7, 6, 10, 19, 14, 18},
                              {0, 3, 2, 1, 11, 10, 9, 8},   {4, 5, 6, 7, 12, 13, 14, 15}};
const int kTetFacets[4][6] = {{0, 3, 2, 7, 9, 6}, {0, 1, 3, 4, 8, 7}, {0, 2, 1, 6, 5, 4}, {1, 2, 3, 5, 9, 8}};

__device__ inline void eig_sym3(double a00, double a11, double a22, double a12, double a02, double a01,
                                double ev[3])
{
    // cyclic Jacobi rotations (each zeroes one off-diagonal entry) until the
    // off-diagonal part vanishes to rounding; eigenvalues sorted descending
    double a[3][3] = {{a00, a01, a02}, {a01, a11, a12}, {a02, a12, a22}};
    for (int sweep = 0; sweep < 16; sweep++) {
        const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        const double dia = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
        if (off == 0.0 || off <= 1e-18 * dia) break;
        for (int pq = 0; pq < 3; pq++) {
            const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2, r = 3 - p - q;
            const double apq = a[p][q];
            if (apq == 0.0) continue;
            const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
            const double t = fabs(theta) > 1e150 ? 0.5 / theta
                                                 : copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
            const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
            a[p][p] -= t * apq;
            a[q][q] += t * apq;
            a[p][q] = a[q][p] = 0.0;
            const double arp = a[r][p], arq = a[r][q];
            a[r][p] = a[p][r] = c * arp - s * arq;
            a[r][q] = a[q][r] = s * arp + c * arq;
        }
    }
    double x = a[0][0], y = a[1][1], z = a[2][2], tmp;
    if (x < y) { tmp = x; x = y; y = tmp; }
    if (y < z) { tmp = y; y = z; z = tmp; }
    if (x < y) { tmp = x; x = y; y = tmp; }
    ev[0] = x;
    ev[1] = y;
    ev[2] = z;
}

__global__ void __launch_bounds__(256) ipvals_kernel(int64_t nip, int nq, const double *__restrict__ st,
                                                      const int32_t *__restrict__ emat,
                                                      const AmrMat *__restrict__ mats, double *__restrict__ out)
{
    const int64_t ip = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (ip >= nip) return;
    double s[6], e[6];
    for (int c = 0; c < 6; c++) {
        s[c] = st[c * nip + ip];
        e[c] = st[(6 + c) * nip + ip];
    }
    const double epa = st[12 * nip + ip];
    const int kind = mats[emat[ip / nq]].kind;
    const double t23 = s[3] / SR2, t13 = s[4] / SR2, t12 = s[5] / SR2;
    const double d1 = s[0] - s[1], d2 = s[1] - s[2], d3 = s[2] - s[0];
    const double j2 = 1.0 / 6.0 * (d1 * d1 + d2 * d2 + d3 * d3) + t23 * t23 + t13 * t13 + t12 * t12;
    double ev[3];
    eig_sym3(s[0], s[1], s[2], t23, t13, t12, ev);
    double *o = out + ip * NF;
    o[0] = s[0];
    o[1] = s[1];
    o[2] = s[2];
    o[3] = t23;
    o[4] = t13;
    o[5] = t12;
    o[6] = sqrt(3.0 * j2);
    o[7] = ev[0];
    o[8] = ev[1];
    o[9] = ev[2];
    o[10] = e[0];
    o[11] = e[1];
    o[12] = e[2];
    o[13] = e[3] / SR2;
    o[14] = e[4] / SR2;
    o[15] = e[5] / SR2;
    o[16] = kind != AMR_MAT_LE ? epa : 0.0;
    o[17] = kind == AMR_MAT_DP ? s[0] + s[1] + s[2] : 0.0;
    o[18] = kind == AMR_MAT_DP ? sqrt(j2) : 0.0;
}

__device__ inline int nterms_of(int m) { return m >= 7 ? 7 : m >= 4 ? 4 : 1; }

// terms 1, x, y, z, xy, yz, xz of the centred, scaled coordinates
__device__ inline void reg_terms(double x, double y, double z, double *t)
{
    t[0] = 1.0;
    t[1] = x;
    t[2] = y;
    t[3] = z;
    t[4] = x * y;
    t[5] = y * z;
    t[6] = x * z;
}

// Sum over the 64 lanes by an xor butterfly: every lane ends with the same
// bits (each step adds the same two operands, in either order).
__device__ inline double wave_sum(double x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// One 64-lane wave per patch.  Lane l holds rows l, l+64, l+128, l+192 of
// the patch's matrix [basis | IP values] (m x NCOL) in registers; each
// Householder step reduces the reflector's norm and its products with the
// remaining columns across the wave.  R and Q^T W (rows 0..nt-1) go through
// LDS to the back substitution, one lane per field.
__global__ void __launch_bounds__(64) patch_fit_kernel(int nq, const int32_t *__restrict__ pptr,
                                                       const int32_t *__restrict__ pelem,
                                                       const double *__restrict__ pcen,
                                                       const double *__restrict__ ipx,
                                                       const double *__restrict__ vals, double *__restrict__ coef,
                                                       int32_t *__restrict__ bad)
{
    constexpr int RPL = MAXM / 64;
    __shared__ double Rs[NT][NCOL];
    const int p = blockIdx.x, lane = threadIdx.x;
    const int e0 = pptr[p];
    const int m = (pptr[p + 1] - e0) * nq;
    const int nt = nterms_of(m);
    const double cx = pcen[p * 4], cy = pcen[p * 4 + 1], cz = pcen[p * 4 + 2], ih = 1.0 / pcen[p * 4 + 3];
    double a[RPL][NCOL];
#pragma unroll
    for (int r = 0; r < RPL; r++) {
        const int i = lane + 64 * r;
#pragma unroll
        for (int c = 0; c < NCOL; c++) a[r][c] = 0.0;
        if (i < m) {
            const int64_t ip = (int64_t)pelem[e0 + i / nq] * nq + i % nq;
            reg_terms((ipx[ip * 3] - cx) * ih, (ipx[ip * 3 + 1] - cy) * ih, (ipx[ip * 3 + 2] - cz) * ih, a[r]);
#pragma unroll
            for (int c = 1; c < NT; c++)
                if (c >= nt) a[r][c] = 0.0;
#pragma unroll
            for (int f = 0; f < NF; f++) a[r][NT + f] = vals[ip * NF + f];
        }
    }
    double rmax = 0.0;
    bool deficient = false;
#pragma unroll
    for (int k = 0; k < NT; k++) {
        if (k >= nt) break;
        // reflector of column k over rows k..m-1 (rows past m hold zeros)
        double v[RPL];
#pragma unroll
        for (int r = 0; r < RPL; r++) v[r] = (r > 0 || lane > k) ? a[r][k] : 0.0;
        double part = 0.0;
#pragma unroll
        for (int r = 0; r < RPL; r++) part += v[r] * v[r];
        const double akk = __shfl(a[0][k], k, 64);
        const double below = wave_sum(part);   // squares of column k below row k
        const double alpha = -copysign(sqrt(below + akk * akk), akk);
        const double vk = akk - alpha;
        const double vv = below + vk * vk;      // v'v of the reflector as applied
        rmax = fmax(rmax, fabs(alpha));
        if (!(fabs(alpha) > 1e-13 * rmax) || vv == 0.0) deficient = true;
        if (lane == k) v[0] = vk;
#pragma unroll
        for (int c = k + 1; c < NCOL; c++) {
            if (c >= nt && c < NT) continue;
            double d = 0.0;
#pragma unroll
            for (int r = 0; r < RPL; r++) d += v[r] * a[r][c];
            const double g = vv != 0.0 ? 2.0 * wave_sum(d) / vv : 0.0;
#pragma unroll
            for (int r = 0; r < RPL; r++) a[r][c] -= g * v[r];
        }
        if (lane == k) a[0][k] = alpha;
    }
    if (lane < nt) {
#pragma unroll
        for (int c = 0; c < NCOL; c++) Rs[lane][c] = a[0][c];
    }
    __syncthreads();
    if (deficient) {
        // rank-deficient basis (small subpatches of a field some elements
        // carry): pinv's minimum-norm fit, which depends on the basis, so it
        // is formed as the reference forms it, on the raw coordinates.
        // One-sided Jacobi SVD of M's columns (rows over the wave), singular
        // values at or below eps * min(m, nt) * smax dropped (Julia's pinv
        // tolerance); the coefficients are then written for the centred,
        // scaled basis node_eval_kernel evaluates (an exact re-expansion).
        __shared__ double Mr[MAXM][NT], Vr[NT][NT];
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            const int i = lane + 64 * r;
            double t[NT] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            if (i < m) {
                const int64_t ip = (int64_t)pelem[e0 + i / nq] * nq + i % nq;
                reg_terms(ipx[ip * 3], ipx[ip * 3 + 1], ipx[ip * 3 + 2], t);
            }
#pragma unroll
            for (int c = 0; c < NT; c++) Mr[i][c] = (i < m && c < nt) ? t[c] : 0.0;
        }
        if (lane < NT)
            for (int c = 0; c < NT; c++) Vr[lane][c] = lane == c ? 1.0 : 0.0;
        __syncthreads();
        for (int sweep = 0; sweep < 60; sweep++) {
            bool rotated = false;
            for (int ca = 0; ca < nt; ca++)
                for (int cb = ca + 1; cb < nt; cb++) {
                    double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
                    for (int r = 0; r < RPL; r++) {
                        const int i = lane + 64 * r;
                        al += Mr[i][ca] * Mr[i][ca];
                        be += Mr[i][cb] * Mr[i][cb];
                        ga += Mr[i][ca] * Mr[i][cb];
                    }
                    al = wave_sum(al);
                    be = wave_sum(be);
                    ga = wave_sum(ga);
                    if (!(fabs(ga) > 1e-15 * sqrt(al * be))) continue;
                    rotated = true;
                    const double zeta = (be - al) / (2.0 * ga);
                    const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                    const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
#pragma unroll
                    for (int r = 0; r < RPL; r++) {
                        const int i = lane + 64 * r;
                        const double ua = Mr[i][ca], ub = Mr[i][cb];
                        Mr[i][ca] = c * ua - sn * ub;
                        Mr[i][cb] = sn * ua + c * ub;
                    }
                    if (lane < NT) {
                        const double va = Vr[lane][ca], vb = Vr[lane][cb];
                        Vr[lane][ca] = c * va - sn * vb;
                        Vr[lane][cb] = sn * va + c * vb;
                    }
                }
            if (!rotated) break;
        }
        __syncthreads();
        double s2[NT], smax = 0.0;
#pragma unroll
        for (int j = 0; j < NT; j++) {
            double n2 = 0.0;
#pragma unroll
            for (int r = 0; r < RPL; r++) n2 += Mr[lane + 64 * r][j] * Mr[lane + 64 * r][j];
            s2[j] = j < nt ? wave_sum(n2) : 0.0;
            smax = fmax(smax, sqrt(s2[j]));
        }
        const double tol = 2.220446049250313e-16 * (double)(m < nt ? m : nt) * smax;
        const double cx = pcen[p * 4], cy = pcen[p * 4 + 1], cz = pcen[p * 4 + 2], h = pcen[p * 4 + 3];
        for (int f = 0; f < NF; f++) {
            // y_j = (M_j . w) / s_j^2 (M_j = U_j s_j), a = V y
            double y[NT];
#pragma unroll
            for (int j = 0; j < NT; j++) {
                double d = 0.0;
#pragma unroll
                for (int r = 0; r < RPL; r++) {
                    const int i = lane + 64 * r;
                    if (i < m) {
                        const int64_t ip = (int64_t)pelem[e0 + i / nq] * nq + i % nq;
                        d += Mr[i][j] * vals[ip * NF + f];
                    }
                }
                d = wave_sum(d);
                y[j] = (j < nt && sqrt(s2[j]) > tol) ? d / s2[j] : 0.0;
            }
            double a[NT];
#pragma unroll
            for (int t = 0; t < NT; t++) {
                double v = 0.0;
#pragma unroll
                for (int j = 0; j < NT; j++) v += Vr[t][j] * y[j];
                a[t] = t < nt ? v : 0.0;
            }
            // x = c + h (xi) etc.: the same polynomial in the centred, scaled terms
            const double b[NT] = {a[0] + a[1] * cx + a[2] * cy + a[3] * cz + a[4] * cx * cy + a[5] * cy * cz +
                                      a[6] * cx * cz,
                                  h * (a[1] + a[4] * cy + a[6] * cz), h * (a[2] + a[4] * cx + a[5] * cz),
                                  h * (a[3] + a[5] * cy + a[6] * cx), h * h * a[4], h * h * a[5], h * h * a[6]};
            if (lane == 0) {
                double *cf = coef + (size_t)p * NT * NF + f;
                for (int t = 0; t < NT; t++) cf[t * NF] = b[t];
            }
        }
        if (lane == 0 && !(smax > 0.0)) bad[p] = 1;
        return;
    }
    if (lane < NF) {
        double x[NT];
#pragma unroll
        for (int t = NT - 1; t >= 0; t--) {
            if (t >= nt) {
                x[t] = 0.0;
                continue;
            }
            double b = Rs[t][NT + lane];
#pragma unroll
            for (int u = t + 1; u < NT; u++)
                if (u < nt) b -= Rs[t][u] * x[u];
            x[t] = b / Rs[t][t];
        }
        double *c = coef + (size_t)p * NT * NF + lane;
#pragma unroll
        for (int t = 0; t < NT; t++) c[t * NF] = x[t];
    }
}

// One thread per (node, field): the field's value from each patch reaching
// the node, summed in patch order, divided by the number of patches.
__global__ void __launch_bounds__(256) node_eval_kernel(int64_t nnodes, int nq, const double *__restrict__ X,
                                                         const int32_t *__restrict__ nptr0,
                                                         const int32_t *__restrict__ npat0,
                                                         const int32_t *__restrict__ nptrx,
                                                         const int32_t *__restrict__ npatx, int fsplit,
                                                         const int32_t *__restrict__ pptr,
                                                         const double *__restrict__ pcen,
                                                         const double *__restrict__ coef, double *__restrict__ out)
{
    const int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (id >= nnodes * NF) return;
    const int64_t n = id / NF;
    const int f = (int)(id - n * NF);
    const int32_t *nptr = f < fsplit ? nptr0 : nptrx, *npat = f < fsplit ? npat0 : npatx;
    double acc = 0.0;
    const int j0 = nptr[n], j1 = nptr[n + 1];
    for (int j = j0; j < j1; j++) {
        const int p = npat[j];
        const int nt = nterms_of((pptr[p + 1] - pptr[p]) * nq);
        const double ih = 1.0 / pcen[p * 4 + 3];
        double t[NT];
        reg_terms((X[n * 3] - pcen[p * 4]) * ih, (X[n * 3 + 1] - pcen[p * 4 + 1]) * ih,
                  (X[n * 3 + 2] - pcen[p * 4 + 2]) * ih, t);
        const double *c = coef + (size_t)p * NT * NF + f;
        double v = t[0] * c[0];
        for (int u = 1; u < nt; u++) v += t[u] * c[u * NF];
        acc += v;
    }
    const double r = acc / (double)(j1 - j0);
    out[id] = r == r ? r : 0.0;   // no patch (0/0) or NaN values: 0, as the reference
}

template <typename T>
int alloc(amr_model *m, T **p, size_t n, int64_t &bytes)
{
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        amr_set_error(std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) + " bytes failed: " +
                      hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? AMR_ERR_NOMEM : AMR_ERR_HIP;
    }
    bytes += (int64_t)(n * sizeof(T));
    m->dev_bytes += (int64_t)(n * sizeof(T));
    return AMR_OK;
}

#define TRY(x)               \
    do {                     \
        int _rc = (x);       \
        if (_rc) return _rc; \
    } while (0)

}  // namespace

struct AmrRecovery {
    int32_t npatch = 0, maxm = 0;   // fit units: patches, then subpatches of the extra fields
    int32_t fsplit = NF;            // fields below it use nptr/npat, the rest nptrx/npatx
    int64_t bytes = 0;
    std::vector<int32_t> reps;   // patches reaching each node
    int32_t *pptr = nullptr, *pelem = nullptr, *nptr = nullptr, *npat = nullptr, *bad = nullptr;
    int32_t *nptrx = nullptr, *npatx = nullptr;
    double *pcen = nullptr, *ipx = nullptr, *vals = nullptr, *coef = nullptr, *nout = nullptr;
};

void amr_recovery_free(amr_model *m)
{
    AmrRecovery *r = m->rec;
    if (!r) return;
    void *ptrs[] = {r->pptr, r->pelem, r->nptr, r->npat, r->bad,  r->pcen,
                    r->ipx,  r->vals,  r->coef, r->nout, r->nptrx, r->npatx};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    m->dev_bytes -= r->bytes;
    delete r;
    m->rec = nullptr;
}

// IP values of the current or committed states into r->vals.
static int launch_ipvals(amr_model *m, AmrRecovery *r, int committed)
{
    if (m->nip_total == 0) return AMR_OK;
    const unsigned g = (unsigned)((m->nip_total + 255) / 256);
    hipLaunchKernelGGL(ipvals_kernel, dim3(g), dim3(256), 0, m->stream, m->nip_total, m->nips,
                       committed ? m->st_com : m->st_cur, m->emat, m->mats, r->vals);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

// The patches of fe-model.jl:522-578, their nodes and the IP coordinates
// (element.jl:160-163), built on the host once per model.
static int build_recovery(amr_model *m)
{
    const int nn = m->nn, nq = m->nips;
    const int64_t ne = m->ne, nnodes = m->nnodes;
    std::vector<int32_t> conn((size_t)ne * nn);
    if (ne > 0) {
        AMR_HIP(hipMemcpyAsync(conn.data(), m->conn, conn.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                               m->stream));
        AMR_HIP(hipStreamSynchronize(m->stream));
    }
    // outer facets (seen once, keyed by their sorted corners) mark boundary nodes
    const bool tet = m->etype == AMR_TET10;
    const int nfac = tet ? 4 : 6, nfn = tet ? 6 : (m->etype == AMR_HEX20 ? 8 : 4), ncf = tet ? 3 : 4;
    const int ncorner = tet ? 4 : 8;
    struct Fk {
        int32_t k[4];
        int64_t id;
    };
    std::vector<Fk> fk((size_t)ne * nfac);
    for (int64_t e = 0; e < ne; e++)
        for (int f = 0; f < nfac; f++) {
            Fk &x = fk[(size_t)e * nfac + f];
            for (int c = 0; c < 4; c++)
                x.k[c] = c < ncf ? conn[(size_t)e * nn + (tet ? kTetFacets[f][c] : kHexFacets[f][c])] : -1;
            std::sort(x.k, x.k + ncf);
            x.id = e * nfac + f;
        }
    std::sort(fk.begin(), fk.end(), [](const Fk &a, const Fk &b) {
        return std::lexicographical_compare(a.k, a.k + 4, b.k, b.k + 4) ||
               (std::equal(a.k, a.k + 4, b.k) && a.id < b.id);
    });
    std::vector<uint8_t> at_bound(nnodes, 0);
    for (size_t i = 0; i < fk.size();) {
        size_t j = i + 1;
        while (j < fk.size() && std::equal(fk[i].k, fk[i].k + 4, fk[j].k)) j++;
        if (j == i + 1) {
            const int64_t e = fk[i].id / nfac;
            const int f = (int)(fk[i].id % nfac);
            for (int c = 0; c < nfn; c++)
                at_bound[conn[(size_t)e * nn + (tet ? kTetFacets[f][c] : kHexFacets[f][c])]] = 1;
        }
        i = j;
    }
    std::vector<std::vector<int32_t>> inner(nnodes), bry(nnodes);
    for (int64_t e = 0; e < ne; e++)
        for (int a = 0; a < ncorner; a++) {
            const int32_t n = conn[(size_t)e * nn + a];
            (at_bound[n] ? bry : inner)[n].push_back((int32_t)e);
        }
    std::vector<uint8_t> has(nnodes, 0);
    for (int64_t n = 0; n < nnodes; n++)
        for (int32_t e : inner[n])
            for (int a = 0; a < nn; a++) has[conn[(size_t)e * nn + a]] = 1;
    std::vector<int64_t> orphans;
    for (int64_t n = 0; n < nnodes; n++)
        if (!has[n] && at_bound[n]) orphans.push_back(n);
    for (int nmin = 3; nmin >= 1 && !orphans.empty(); nmin--) {
        for (int64_t n : orphans)
            if ((int)bry[n].size() >= nmin) inner[n] = bry[n];
        for (int64_t n : orphans)
            for (int32_t e : inner[n])
                for (int a = 0; a < nn; a++) has[conn[(size_t)e * nn + a]] = 1;
        std::vector<int64_t> left;
        for (int64_t n : orphans)
            if (!has[n]) left.push_back(n);
        orphans.swap(left);
    }
    // IP coordinates C'*N
    std::vector<double> ipx((size_t)m->nip_total * 3);
    for (int64_t e = 0; e < ne; e++)
        for (int q = 0; q < nq; q++)
            for (int k = 0; k < 3; k++) {
                double x = 0.0;
                for (int a = 0; a < nn; a++)
                    x += m->Xh[(size_t)conn[(size_t)e * nn + a] * 3 + k] * m->Nip[(size_t)q * nn + a];
                ipx[((size_t)e * nq + q) * 3 + k] = x;
            }
    // patches in node order; their nodes in order of first appearance
    std::vector<int32_t> pptr{0}, pelem, owner, pn_node, pn_patch;
    std::vector<double> pcen;
    std::vector<int32_t> seen(nnodes, -1);
    int32_t maxm = 0;
    for (int64_t n = 0; n < nnodes; n++) {
        if (inner[n].empty()) continue;
        const int32_t p = (int32_t)owner.size();
        const int m_rows = (int)inner[n].size() * nq;
        if (m_rows > MAXM) {
            amr_set_error("nodal recovery: the patch of node " + std::to_string(n + 1) + " has " +
                          std::to_string(m_rows) + " integration points (at most " + std::to_string(MAXM) + ")");
            return AMR_ERR_ARG;
        }
        maxm = std::max(maxm, m_rows);
        owner.push_back((int32_t)n);
        const double c[3] = {m->Xh[n * 3], m->Xh[n * 3 + 1], m->Xh[n * 3 + 2]};
        double h = 0.0;
        for (int32_t e : inner[n]) {
            pelem.push_back(e);
            for (int q = 0; q < nq; q++)
                for (int k = 0; k < 3; k++) h = std::max(h, std::fabs(ipx[((size_t)e * nq + q) * 3 + k] - c[k]));
            for (int a = 0; a < nn; a++) {
                const int32_t v = conn[(size_t)e * nn + a];
                if (seen[v] == p) continue;
                seen[v] = p;
                pn_node.push_back(v);
                for (int k = 0; k < 3; k++) h = std::max(h, std::fabs(m->Xh[(size_t)v * 3 + k] - c[k]));
            }
        }
        pptr.push_back((int32_t)pelem.size());
        pcen.insert(pcen.end(), {c[0], c[1], c[2], h > 0.0 ? h : 1.0});
        pn_patch.resize(pn_node.size(), p);
    }
    // a model that mixes material kinds: the subpatches of the elements that
    // carry the extra fields (those of its one plastic kind), as fit units
    // after the patches, with their nodes in order of first appearance
    const int32_t npatch0 = (int32_t)owner.size();
    int32_t fsplit = NF;
    std::vector<int32_t> px_node, px_unit;
    if (m->elem_kind == -2) {
        int plastic = -1;
        for (int8_t k : m->ekind)
            if (k != AMR_MAT_LE) {
                if (plastic >= 0 && k != plastic) {
                    amr_set_error("amr_nodal_recovery: von Mises and Drucker-Prager elements in one model (their "
                                  "extra output fields share columns)");
                    return AMR_ERR_ARG;
                }
                plastic = k;
            }
        fsplit = 16;   // the stress_strain_dict fields, carried by every element
        std::fill(seen.begin(), seen.end(), -1);
        for (int32_t p = 0; p < npatch0 && plastic >= 0; p++) {
            const int32_t u = (int32_t)pptr.size() - 1;
            const size_t before = pelem.size();
            for (int32_t i = pptr[p]; i < pptr[p + 1]; i++) {
                const int32_t e = pelem[i];
                if (m->ekind[e] == plastic) pelem.push_back(e);
            }
            if (pelem.size() == before) continue;
            pptr.push_back((int32_t)pelem.size());
            for (int k = 0; k < 4; k++) pcen.push_back(pcen[(size_t)p * 4 + k]);
            for (size_t i = before; i < pelem.size(); i++)
                for (int a = 0; a < nn; a++) {
                    const int32_t v = conn[(size_t)pelem[i] * nn + a];
                    if (seen[v] == u) continue;
                    seen[v] = u;
                    px_node.push_back(v);
                    px_unit.push_back(u);
                }
        }
    }
    // node -> patches, ascending patch (the reference's summation order)
    AmrRecovery *r = m->rec ? m->rec : (m->rec = new AmrRecovery);   // amr_ip_values may have made it
    r->npatch = (int32_t)pptr.size() - 1;
    r->maxm = maxm;
    r->fsplit = fsplit;
    r->reps.assign(nnodes, 0);
    for (int32_t v : pn_node) r->reps[v]++;
    std::vector<int32_t> nptr(nnodes + 1, 0);
    for (int64_t n = 0; n < nnodes; n++) nptr[n + 1] = nptr[n] + r->reps[n];
    std::vector<int32_t> npat(pn_node.size()), fill(nptr.begin(), nptr.end() - 1);
    for (size_t i = 0; i < pn_node.size(); i++) npat[fill[pn_node[i]]++] = pn_patch[i];
    std::vector<int32_t> nptrx(nnodes + 1, 0), npatx(px_node.size());
    for (int32_t v : px_node) nptrx[v + 1]++;
    for (int64_t n = 0; n < nnodes; n++) nptrx[n + 1] += nptrx[n];
    {
        std::vector<int32_t> fx(nptrx.begin(), nptrx.end() - 1);
        for (size_t i = 0; i < px_node.size(); i++) npatx[fx[px_node[i]]++] = px_unit[i];
    }

    const int np = r->npatch;
    int rc = AMR_OK;
    if (!rc) rc = alloc(m, &r->pptr, pptr.size(), r->bytes);
    if (!rc) rc = alloc(m, &r->pelem, pelem.size(), r->bytes);
    if (!rc) rc = alloc(m, &r->nptr, nptr.size(), r->bytes);
    if (!rc) rc = alloc(m, &r->npat, npat.size(), r->bytes);
    if (!rc) rc = alloc(m, &r->nptrx, nptrx.size(), r->bytes);
    if (!rc) rc = alloc(m, &r->npatx, npatx.size(), r->bytes);
    if (!rc) rc = alloc(m, &r->bad, (size_t)np, r->bytes);
    if (!rc) rc = alloc(m, &r->pcen, pcen.size(), r->bytes);
    if (!rc) rc = alloc(m, &r->ipx, ipx.size(), r->bytes);
    if (!rc && !r->vals) rc = alloc(m, &r->vals, (size_t)m->nip_total * NF, r->bytes);
    if (!rc) rc = alloc(m, &r->coef, (size_t)np * NT * NF, r->bytes);
    if (!rc) rc = alloc(m, &r->nout, (size_t)nnodes * NF, r->bytes);
    if (rc) {
        amr_recovery_free(m);
        return rc;
    }
    hipStream_t s = m->stream;
    struct Up {
        void *d;
        const void *h;
        size_t n;
    } ups[] = {{r->pptr, pptr.data(), pptr.size() * 4},   {r->pelem, pelem.data(), pelem.size() * 4},
               {r->nptr, nptr.data(), nptr.size() * 4},   {r->npat, npat.data(), npat.size() * 4},
               {r->nptrx, nptrx.data(), nptrx.size() * 4}, {r->npatx, npatx.data(), npatx.size() * 4},
               {r->pcen, pcen.data(), pcen.size() * 8},   {r->ipx, ipx.data(), ipx.size() * 8}};
    hipError_t e = hipSuccess;
    for (const Up &u : ups)
        if (u.n && e == hipSuccess) e = hipMemcpyAsync(u.d, u.h, u.n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);   // the host vectors go out of scope
    if (e != hipSuccess) {
        amr_set_error(std::string("HIP error: ") + hipGetErrorString(e) + " uploading the recovery patches");
        amr_recovery_free(m);
        return AMR_ERR_HIP;
    }
    return AMR_OK;
}

extern "C" int amr_ip_values(amr_model *m, int committed, double *out)
{
    if (!m || !out) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    if (m->nip_total == 0) return AMR_OK;
    // the IP values need no patches: only their buffer
    if (!m->rec) m->rec = new AmrRecovery;
    AmrRecovery *r = m->rec;
    if (!r->vals) TRY(alloc(m, &r->vals, (size_t)m->nip_total * NF, r->bytes));
    TRY(launch_ipvals(m, r, committed));
    AMR_HIP(hipMemcpyAsync(out, r->vals, sizeof(double) * NF * m->nip_total, hipMemcpyDeviceToHost, m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    return AMR_OK;
}

extern "C" int amr_nodal_recovery(amr_model *m, int committed, double *out, int32_t *reps)
{
    if (!m || !out) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    if (!m->rec || !m->rec->pptr) TRY(build_recovery(m));
    AmrRecovery *r = m->rec;
    if (reps) std::memcpy(reps, r->reps.data(), sizeof(int32_t) * m->nnodes);
    hipStream_t s = m->stream;
    TRY(launch_ipvals(m, r, committed));
    if (r->npatch > 0) {
        AMR_HIP(hipMemsetAsync(r->bad, 0, sizeof(int32_t) * r->npatch, s));
        hipLaunchKernelGGL(patch_fit_kernel, dim3(r->npatch), dim3(64), 0, s, m->nips, r->pptr, r->pelem,
                           r->pcen, r->ipx, r->vals, r->coef, r->bad);
        AMR_HIP(hipGetLastError());
    }
    const unsigned g = (unsigned)((m->nnodes * NF + 255) / 256);
    hipLaunchKernelGGL(node_eval_kernel, dim3(g), dim3(256), 0, s, m->nnodes, m->nips, m->X, r->nptr, r->npat,
                       r->nptrx, r->npatx, r->fsplit,
                       r->pptr, r->pcen, r->coef, r->nout);
    AMR_HIP(hipGetLastError());
    AMR_HIP(hipMemcpyAsync(out, r->nout, sizeof(double) * NF * m->nnodes, hipMemcpyDeviceToHost, s));
    std::vector<int32_t> bad(r->npatch);
    if (r->npatch) AMR_HIP(hipMemcpyAsync(bad.data(), r->bad, sizeof(int32_t) * r->npatch, hipMemcpyDeviceToHost, s));
    AMR_HIP(hipStreamSynchronize(s));
    for (int32_t p = 0; p < r->npatch; p++)
        if (bad[p]) {
            amr_set_error("amr_nodal_recovery: rank-deficient least-squares basis in patch " + std::to_string(p));
            return AMR_ERR_SOLVE;
        }
    return AMR_OK;
}
