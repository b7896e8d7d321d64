// Element-level HIP kernels for gfx950: element stiffness, deterministic
// assembly gather, per-IP constitutive update and internal-force gather.
//
// Floating-point order.  Every kernel here evaluates its sums in one fixed
// order: dot products are fma chains in ascending index order starting from
// 0, K += coef*t and dF += coef*t are a single fma, divisions stay divisions,
// and the library is compiled with -ffp-contract=off.  Structural zeros of
// the Mandel B matrix are skipped; an fma with a zero factor leaves a
// non-zero accumulator unchanged, so skipping them does not change a bit.
#include <cmath>
#include <cstring>
#include <map>

#include <mutex>

#include "amr_internal.h"

namespace {

constexpr double SR2 = 1.4142135623730951;  // sqrt(2.0)
constexpr double DEPS = 2.220446049250313e-16;  // eps(Float64)

// ------------------------------------------------------------- materials

// Elastic Mandel stiffness, src/mech/mat/linear-elastic.jl:110-117.
__device__ __host__ inline void calc_De(double E, double nu, double *D)
{
    double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double a = c * (1.0 - nu), b = c * nu, g = c * (1.0 - 2.0 * nu);
    for (int i = 0; i < 36; i++) D[i] = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) D[i * 6 + j] = (i == j) ? a : b;
    for (int i = 3; i < 6; i++) D[i * 6 + i] = g;
}

__device__ inline void mat6_vec(const double *A, const double *x, double *y)
{
    for (int i = 0; i < 6; i++) {
        double acc = 0.0;
        for (int j = 0; j < 6; j++) acc = fma(A[i * 6 + j], x[j], acc);
        y[i] = acc;
    }
}

__device__ inline double dot6(const double *a, const double *b)
{
    double acc = 0.0;
    for (int i = 0; i < 6; i++) acc = fma(a[i], b[i], acc);
    return acc;
}

// deviatoric part via the 6x6 projector product (src/tools/tensors.jl:10-18)
__device__ inline void deviator(const double *x, double *s)
{
    const double d = 2.0 / 3.0, o = -1.0 / 3.0;
    s[0] = fma(o, x[2], fma(o, x[1], d * x[0]));
    s[1] = fma(o, x[2], fma(d, x[1], o * x[0]));
    s[2] = fma(d, x[2], fma(o, x[1], o * x[0]));
    s[3] = x[3];
    s[4] = x[4];
    s[5] = x[5];
}

// J2 of a Mandel vector (src/tools/tensors.jl:24-27)
__device__ inline double j2_of(const double *x)
{
    double t23 = x[3] / SR2, t13 = x[4] / SR2, t12 = x[5] / SR2;
    double a = x[0] - x[1], b = x[1] - x[2], c = x[2] - x[0];
    return (1.0 / 6.0) * (a * a + b * b + c * c) + t23 * t23 + t13 * t13 + t12 * t12;
}

// calcD for LE / VM 3D / DP (src/mech/mat/linear-elastic.jl:125-127,
// von-mises.jl:112-125, drucker-prager.jl:86-109).  st: sig[6] .. dlam.
__device__ void calc_D(const AmrMat &m, const double *st, double *D)
{
    calc_De(m.p[0], m.p[1], D);
    if (m.kind == AMR_MAT_LE || st[13] == 0.0) return;
    double De[36];
    for (int i = 0; i < 36; i++) De[i] = D[i];
    if (m.kind == AMR_MAT_VM) {
        double s[6], n[6], v[6];
        deviator(st, s);
        double nrm = sqrt(dot6(s, s));
        for (int i = 0; i < 6; i++) n[i] = (sqrt(1.5) * s[i]) / nrm;
        mat6_vec(De, n, v);
        double den = dot6(n, v) - sqrt(1.5) * (-m.p[3]);
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) D[i * 6 + j] = De[i * 6 + j] - (v[i] * v[j]) / den;
    } else {
        double alpha = m.p[2], H = m.p[4];
        double Nu[6], V[6], a[6], b[6];
        const double I2[6] = {1, 1, 1, 0, 0, 0};
        if (j2_of(st) != 0.0) {
            double s[6];
            deviator(st, s);
            double ns = sqrt(dot6(s, s));
            for (int i = 0; i < 6; i++) V[i] = alpha * I2[i] + (s[i] / ns) / SR2;
            double nv = sqrt(dot6(V, V));
            for (int i = 0; i < 6; i++) Nu[i] = V[i] / nv;
        } else {
            for (int i = 0; i < 6; i++) Nu[i] = V[i] = (1.0 / sqrt(3.0)) * I2[i];
        }
        mat6_vec(De, Nu, a);
        mat6_vec(De, V, b);
        double den = dot6(b, Nu) + H;
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) D[i * 6 + j] = De[i * 6 + j] - (a[i] * b[j]) / den;
    }
}

// update_state! for LE / VM 3D / DP (src/mech/mat/linear-elastic.jl:130-137,
// von-mises.jl:128-156, drucker-prager.jl:112-149).  st is updated in place;
// returns 0, or 1 on a material failure (VM "negative sqrt(J2D)").
__device__ int update_ip(const AmrMat &m, double *st, const double *de, double *ds)
{
    double De[36], sini[6], str[6], t[6];
    calc_De(m.p[0], m.p[1], De);
    for (int i = 0; i < 6; i++) sini[i] = st[i];
    mat6_vec(De, de, t);
    for (int i = 0; i < 6; i++) str[i] = st[i] + t[i];
    if (m.kind == AMR_MAT_LE) {
        for (int i = 0; i < 6; i++) {
            st[6 + i] += de[i];
            st[i] = str[i];
            ds[i] = t[i];
        }
        return 0;
    }
    if (m.kind == AMR_MAT_VM) {
        double E = m.p[0], nu = m.p[1], fy = m.p[2], H = m.p[3];
        double j2 = j2_of(str);
        double ftr = sqrt(3.0 * j2) - fy - H * st[12];
        if (ftr < 1e-8) {
            st[13] = 0.0;
            for (int i = 0; i < 6; i++) st[i] = str[i];
        } else {
            double G = E / (2.0 * (1.0 + nu));
            double dl = ftr / (3.0 * G + sqrt(1.5) * H);
            st[13] = dl;
            double sj = sqrt(j2);
            if (!(sj - dl * sqrt(3.0) * G >= 0.0)) return 1;
            double f = 1.0 - sqrt(3.0) * G * dl / sj;
            double s[6];
            deviator(str, s);
            for (int i = 0; i < 6; i++) s[i] = f * s[i];
            double ns = sqrt(dot6(s, s));
            double c = sqrt(6.0) * G * dl;
            for (int i = 0; i < 6; i++) st[i] = str[i] - (c * s[i]) / ns;
            st[12] += dl;
        }
    } else {
        double E = m.p[0], nu = m.p[1], alpha = m.p[2], kappa = m.p[3], H = m.p[4];
        double j1 = str[0] + str[1] + str[2];
        double j2 = j2_of(str);
        double ftr = alpha * j1 + sqrt(j2) - kappa - H * st[12];
        if (ftr < 1e-8) {
            st[13] = 0.0;
            for (int i = 0; i < 6; i++) st[i] = str[i];
        } else {
            double K = E / (3.0 * (1.0 - 2.0 * nu)), G = E / (2.0 * (1.0 + nu));
            double n = 1.0 / sqrt(3.0 * alpha * alpha + 0.5);
            double sj = sqrt(j2);
            if (sj - st[13] * n * G > 0.0) {  // previous dgamma, as the reference
                double dg = ftr / (9.0 * alpha * alpha * n * K + n * G + H);
                st[13] = dg;
                double j1n = j1 - 9.0 * dg * alpha * n * K;
                double mm = 1.0 - dg * n * G / sj;
                double s[6];
                deviator(str, s);
                for (int i = 0; i < 6; i++) st[i] = mm * s[i] + (i < 3 ? j1n / 3.0 : 0.0);
            } else {
                double dg = (alpha * j1 - kappa - H * st[12]) / (3.0 * sqrt(3.0) * alpha * K + H);
                st[13] = dg;
                double j1n = j1 - 3.0 * sqrt(3.0) * dg * K;
                for (int i = 0; i < 6; i++) st[i] = (i < 3 ? j1n / 3.0 : 0.0);
            }
            st[12] += st[13];
        }
    }
    for (int i = 0; i < 6; i++) {
        st[6 + i] += de[i];
        ds[i] = st[i] - sini[i];
    }
    return 0;
}

// Jacobian inverse by cofactors; returns det.  Same expressions as the
// oracle's ip_geometry.
__device__ inline double inv3(const double *J, double *inv)
{
    double c00 = J[4] * J[8] - J[5] * J[7];
    double c01 = J[5] * J[6] - J[3] * J[8];
    double c02 = J[3] * J[7] - J[4] * J[6];
    double det = J[0] * c00 + J[1] * c01 + J[2] * c02;
    inv[0] = c00 / det;
    inv[3] = c01 / det;
    inv[6] = c02 / det;
    inv[1] = (J[2] * J[7] - J[1] * J[8]) / det;
    inv[4] = (J[0] * J[8] - J[2] * J[6]) / det;
    inv[7] = (J[1] * J[6] - J[0] * J[7]) / det;
    inv[2] = (J[1] * J[5] - J[2] * J[4]) / det;
    inv[5] = (J[2] * J[3] - J[0] * J[5]) / det;
    inv[8] = (J[0] * J[4] - J[1] * J[3]) / det;
    return det;
}

__device__ inline double inv3_entry(const double *J, int k, double *det_out)
{
    double c00 = J[4] * J[8] - J[5] * J[7];
    double c01 = J[5] * J[6] - J[3] * J[8];
    double c02 = J[3] * J[7] - J[4] * J[6];
    double det = J[0] * c00 + J[1] * c01 + J[2] * c02;
    *det_out = det;
    double num;
    switch (k) {
    case 0: num = c00; break;
    case 3: num = c01; break;
    case 6: num = c02; break;
    case 1: num = J[2] * J[7] - J[1] * J[8]; break;
    case 4: num = J[0] * J[8] - J[2] * J[6]; break;
    case 7: num = J[1] * J[6] - J[0] * J[7]; break;
    case 2: num = J[1] * J[5] - J[2] * J[4]; break;
    case 5: num = J[2] * J[3] - J[0] * J[5]; break;
    default: num = J[0] * J[4] - J[1] * J[3]; break;
    }
    return num / det;
}

// ---------------------------------------------------- element stiffness
//
// One 64-lane workgroup per element.  Lane (ch, b) = lane ch*NN + b owns the
// node blocks (a, b) for a in chunk ch (ACH consecutive row nodes) and keeps
// their 9*ACH entries in registers across all IPs.  Per IP: 9 lanes build J
// and its inverse, NN lanes build dN/dX (and /sqrt2) into LDS, one lane
// evaluates calcD for plastic states, then every lane forms DB for its
// column node (54 fma) and updates its blocks (36 fma per block).

// Each kernel's body is a device function of its workgroup index, inlined
// into the single-model kernel (the workgroup index is blockIdx.x) and into
// the batched kernel of several models (the index within its model).
template <int NN, int NQ, int ACH>
__device__ __attribute__((always_inline)) inline void
ke_kernel_at(int64_t bid_, int64_t ne, const double *__restrict__ X, const int32_t *__restrict__ conn,
             const int32_t *__restrict__ emat, const AmrMat *__restrict__ mats, const double *__restrict__ dNdR,
             const double *__restrict__ wip, const double *__restrict__ st, int64_t nip, int all_elastic,
             double *__restrict__ Ke, int32_t *__restrict__ err)
{
    constexpr int NCH = (NN + ACH - 1) / ACH;
    const int64_t e = bid_;
    const int lane = threadIdx.x;
    __shared__ double sC[NN * 3];
    __shared__ double sJ[9], sInv[9], sCoef;
    __shared__ double sG[NN * 6];
    __shared__ double sD[36];
    __shared__ int sBad;

    for (int i = lane; i < NN * 3; i += 64) sC[i] = X[(int64_t)conn[e * NN + i / 3] * 3 + i % 3];
    const AmrMat m = mats[emat[e]];
    if (lane == 0) {
        sBad = 0;
        if (all_elastic || m.kind == AMR_MAT_LE) calc_De(m.p[0], m.p[1], sD);
    }
    const bool elastic = all_elastic || m.kind == AMR_MAT_LE;
    const int b = lane % NN, ch = lane / NN;
    const bool active = lane < NN * NCH;
    double acc[ACH][9];
#pragma unroll
    for (int i = 0; i < ACH; i++)
#pragma unroll
        for (int s = 0; s < 9; s++) acc[i][s] = 0.0;
    __syncthreads();

    for (int q = 0; q < NQ; q++) {
        const double *dq = dNdR + q * NN * 3;
        if (lane < 9) {
            int i = lane / 3, j = lane % 3;
            double a = 0.0;
            for (int n = 0; n < NN; n++) a = fma(sC[n * 3 + i], dq[n * 3 + j], a);
            sJ[lane] = a;
        }
        __syncthreads();
        if (lane < 9) {
            double det;
            sInv[lane] = inv3_entry(sJ, lane, &det);
            if (lane == 0) {
                sCoef = det * wip[q];
                if (!(det > 0.0)) sBad = 1;
            }
        } else if (lane == 32 && !elastic) {
            double s14[AMR_NSTATE];
            int64_t ip = e * NQ + q;
            for (int f = 0; f < AMR_NSTATE; f++) s14[f] = st[f * nip + ip];
            calc_D(m, s14, sD);
        }
        __syncthreads();
        if (lane < NN) {
            double g[3];
            for (int j = 0; j < 3; j++) {
                double a = 0.0;
                for (int i = 0; i < 3; i++) a = fma(dq[lane * 3 + i], sInv[i * 3 + j], a);
                g[j] = a;
            }
            sG[lane * 6 + 0] = g[0];
            sG[lane * 6 + 1] = g[1];
            sG[lane * 6 + 2] = g[2];
            sG[lane * 6 + 3] = g[0] / SR2;
            sG[lane * 6 + 4] = g[1] / SR2;
            sG[lane * 6 + 5] = g[2] / SR2;
        }
        __syncthreads();
        if (active) {
            const double coef = sCoef;
            const double gx = sG[b * 6], gy = sG[b * 6 + 1], gz = sG[b * 6 + 2];
            const double hx = sG[b * 6 + 3], hy = sG[b * 6 + 4], hz = sG[b * 6 + 5];
            double DB[6][3];
#pragma unroll
            for (int k = 0; k < 6; k++) {
                DB[k][0] = fma(sD[k * 6 + 5], hy, fma(sD[k * 6 + 4], hz, fma(sD[k * 6 + 0], gx, 0.0)));
                DB[k][1] = fma(sD[k * 6 + 5], hx, fma(sD[k * 6 + 3], hz, fma(sD[k * 6 + 1], gy, 0.0)));
                DB[k][2] = fma(sD[k * 6 + 4], hx, fma(sD[k * 6 + 3], hy, fma(sD[k * 6 + 2], gz, 0.0)));
            }
#pragma unroll
            for (int ia = 0; ia < ACH; ia++) {
                const int a = ch * ACH + ia;
                if (a < NN) {
                    const double ax = sG[a * 6], ay = sG[a * 6 + 1], az = sG[a * 6 + 2];
                    const double bx = sG[a * 6 + 3], by = sG[a * 6 + 4], bz = sG[a * 6 + 5];
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        // row r=0 (x): B rows 0 (ax), 4 (bz), 5 (by)
                        double t0 = fma(by, DB[5][c], fma(bz, DB[4][c], fma(ax, DB[0][c], 0.0)));
                        // r=1 (y): rows 1 (ay), 3 (bz), 5 (bx)
                        double t1 = fma(bx, DB[5][c], fma(bz, DB[3][c], fma(ay, DB[1][c], 0.0)));
                        // r=2 (z): rows 2 (az), 3 (by), 4 (bx)
                        double t2 = fma(bx, DB[4][c], fma(by, DB[3][c], fma(az, DB[2][c], 0.0)));
                        acc[ia][0 * 3 + c] = fma(coef, t0, acc[ia][0 * 3 + c]);
                        acc[ia][1 * 3 + c] = fma(coef, t1, acc[ia][1 * 3 + c]);
                        acc[ia][2 * 3 + c] = fma(coef, t2, acc[ia][2 * 3 + c]);
                    }
                }
            }
        }
        __syncthreads();
    }
    if (active) {
#pragma unroll
        for (int ia = 0; ia < ACH; ia++) {
            const int a = ch * ACH + ia;
            if (a < NN) {
                double *o = Ke + ((e * NN + a) * NN + b) * 9;
#pragma unroll
                for (int s = 0; s < 9; s++) o[s] = acc[ia][s];
            }
        }
    }
    if (lane == 0 && sBad) atomicMin(err, (int32_t)(e + 1));
}

template <int NN, int NQ, int ACH>
__global__ void __launch_bounds__(64) ke_kernel(int64_t ne, const double *__restrict__ X,
                                  const int32_t *__restrict__ conn, const int32_t *__restrict__ emat,
                                  const AmrMat *__restrict__ mats, const double *__restrict__ dNdR,
                                  const double *__restrict__ wip, const double *__restrict__ st, int64_t nip,
                                  int all_elastic, double *__restrict__ Ke, int32_t *__restrict__ err)
{
    ke_kernel_at<NN, NQ, ACH>(blockIdx.x, ne, X, conn, emat, mats, dNdR, wip, st, nip, all_elastic, Ke, err);
}

// ---------------------------------- element stiffness on the matrix cores
//
// HEX20 (60 dofs): Ke = sum_q Bq' (coef_q D_q Bq) as one product
// [B1' ... Bnq'] [c1 D1 B1; ...] with FP64 MFMA 16x16x4 tiles: the 60x60
// result padded to 64x64 is 4x4 tiles of 16x16, of which the 10 on or above
// the diagonal are computed and mirrored.  The inner dimension runs over the
// IPs, 8 per IP (the six Mandel rows and two zero rows: two MFMA k-steps).
// One 64-lane workgroup per element.  First the geometry of every IP (J, its
// inverse and coef by lanes q, calcD for plastic states, dN/dX and dN/dX/sqrt2
// per (IP, node)) is staged in LDS, so the element needs three barriers.
// Every product has the exact kernel's operands; only the summation order
// differs (the matrix cores accumulate in their own order), so Ke agrees with
// the exact path to rounding, not bit for bit.
//
// MFMA f64 16x16x4 operands (one value per lane): A[m][k] at lane
// m + 16 k, B[k][n] at lane n + 16 k; the 16x16 result D[l/16 + 4 r][l%16]
// in element r of lane l's four-value accumulator.
typedef double f64x4m __attribute__((ext_vector_type(4)));

// Bq[row][m] (Mandel B, src/mech/elem/mech-solid.jl:106-118) for dof m =
// 3 node + dir from the node's (g, h = g/sqrt2); rows 6, 7 and dofs >= 60 are 0
__device__ inline double b_entry(int row, int m, const double *g, const double *h)
{
    if (m >= 60) return 0.0;
    const int n = m / 3, d = m % 3;
    const double *gn = g + n * 3, *hn = h + n * 3;
    switch (row) {
    case 0: return d == 0 ? gn[0] : 0.0;
    case 1: return d == 1 ? gn[1] : 0.0;
    case 2: return d == 2 ? gn[2] : 0.0;
    case 3: return d == 1 ? hn[2] : d == 2 ? hn[1] : 0.0;
    case 4: return d == 0 ? hn[2] : d == 2 ? hn[0] : 0.0;
    case 5: return d == 0 ? hn[1] : d == 1 ? hn[0] : 0.0;
    }
    return 0.0;
}

// (D Bq)[row][n]: the exact kernel's three-term chains for column n
__device__ inline double db_entry(int row, int n, const double *D, const double *g, const double *h)
{
    if (n >= 60 || row >= 6) return 0.0;
    const int b = n / 3, c = n % 3;
    const double *Dk = D + row * 6;
    const double *gb = g + b * 3, *hb = h + b * 3;
    if (c == 0) return fma(Dk[5], hb[1], fma(Dk[4], hb[2], fma(Dk[0], gb[0], 0.0)));
    if (c == 1) return fma(Dk[5], hb[0], fma(Dk[3], hb[2], fma(Dk[1], gb[1], 0.0)));
    return fma(Dk[4], hb[0], fma(Dk[3], hb[1], fma(Dk[2], gb[2], 0.0)));
}

template <int NQ, int NQC>
__device__ __attribute__((always_inline)) inline void
ke_mfma_hex20_kernel_at(int64_t bid_, int64_t ne, const double *__restrict__ X, const int32_t *__restrict__ conn,
                        const int32_t *__restrict__ emat, const AmrMat *__restrict__ mats,
                        const double *__restrict__ dNdR, const double *__restrict__ wip,
                        const double *__restrict__ st, int64_t nip, int all_elastic, double *__restrict__ Ke,
                        int32_t *__restrict__ err)
{
    constexpr int NN = 20;
    static_assert(NQ % NQC == 0, "IP chunks");
    const int64_t e = bid_;
    const int lane = threadIdx.x;
    __shared__ double sC[NN * 3];
    __shared__ double sG[NQC][NN * 3], sH[NQC][NN * 3];
    __shared__ double sInv[NQC][9], sCoef[NQC];
    __shared__ double sD[NQC][36];
    __shared__ int sBad;
    for (int i = lane; i < NN * 3; i += 64) sC[i] = X[(int64_t)conn[e * NN + i / 3] * 3 + i % 3];
    const AmrMat m = mats[emat[e]];
    const bool elastic = all_elastic || m.kind == AMR_MAT_LE;
    if (lane == 0) {
        sBad = 0;
        if (elastic) calc_De(m.p[0], m.p[1], sD[0]);
    }
    const int lr = lane % 16, lk = lane / 16;
    f64x4m acc[10];
#pragma unroll
    for (int t = 0; t < 10; t++) acc[t] = f64x4m{0.0, 0.0, 0.0, 0.0};
    for (int q0 = 0; q0 < NQ; q0 += NQC) {
        __syncthreads();
        // geometry of the chunk's IPs: lanes q
        if (lane < NQC) {
            const int q = q0 + lane;
            const double *dq = dNdR + q * NN * 3;
            double J[9], inv[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double a = 0.0;
                    for (int n = 0; n < NN; n++) a = fma(sC[n * 3 + i], dq[n * 3 + j], a);
                    J[i * 3 + j] = a;
                }
            const double det = inv3(J, inv);
            if (!(det > 0.0)) sBad = 1;
            for (int k = 0; k < 9; k++) sInv[lane][k] = inv[k];
            sCoef[lane] = det * wip[q];
            if (!elastic) {
                double s14[AMR_NSTATE];
                const int64_t ip = e * NQ + q;
                for (int f = 0; f < AMR_NSTATE; f++) s14[f] = st[f * nip + ip];
                calc_D(m, s14, sD[lane]);
            }
        }
        __syncthreads();
        for (int t = lane; t < NQC * NN; t += 64) {
            const int qc = t / NN, n = t % NN;
            const double *dq = dNdR + (q0 + qc) * NN * 3;
            for (int j = 0; j < 3; j++) {
                double a = 0.0;
                for (int i = 0; i < 3; i++) a = fma(dq[n * 3 + i], sInv[qc][i * 3 + j], a);
                sG[qc][n * 3 + j] = a;
                sH[qc][n * 3 + j] = a / SR2;
            }
        }
        __syncthreads();
        for (int qc = 0; qc < NQC; qc++) {
            const double c = sCoef[qc];
            const double *D = sD[elastic ? 0 : qc];
            const double *g = sG[qc], *h = sH[qc];
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                const int row = ks * 4 + lk;
                double a[4], b[4];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    a[t] = b_entry(row, t * 16 + lr, g, h);
                    b[t] = c * db_entry(row, t * 16 + lr, D, g, h);
                }
                int idx = 0;
#pragma unroll
                for (int mt = 0; mt < 4; mt++)
#pragma unroll
                    for (int nt = mt; nt < 4; nt++) {
                        acc[idx] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mt], b[nt], acc[idx], 0, 0, 0);
                        idx++;
                    }
            }
        }
    }
    // Ke in the node-block layout [a][b][3x3]; the tiles below the diagonal
    // are the mirror of those above
    double *Ke_e = Ke + e * (int64_t)(NN * NN * 9);
    int idx = 0;
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = mt; nt < 4; nt++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = mt * 16 + lk + 4 * r, col = nt * 16 + lr;
                if (row < 60 && col < 60) {
                    const double v = acc[idx][r];
                    Ke_e[((row / 3) * NN + col / 3) * 9 + (row % 3) * 3 + col % 3] = v;
                    if (mt != nt) Ke_e[((col / 3) * NN + row / 3) * 9 + (col % 3) * 3 + row % 3] = v;
                }
            }
            idx++;
        }
    if (lane == 0 && sBad) atomicMin(err, (int32_t)(e + 1));
}

template <int NQ, int NQC>
__global__ void __launch_bounds__(64, 2) ke_mfma_hex20_kernel(int64_t ne, const double *__restrict__ X,
                                  const int32_t *__restrict__ conn, const int32_t *__restrict__ emat,
                                  const AmrMat *__restrict__ mats, const double *__restrict__ dNdR,
                                  const double *__restrict__ wip, const double *__restrict__ st, int64_t nip,
                                  int all_elastic, double *__restrict__ Ke, int32_t *__restrict__ err)
{
    ke_mfma_hex20_kernel_at<NQ, NQC>(blockIdx.x, ne, X, conn, emat, mats, dNdR, wip, st, nip, all_elastic, Ke, err);
}

// Linear elastic elements (every material LinearElastic).  With the Lame
// constants lambda = c nu, mu = c (1 - 2 nu) / 2 of the Mandel De
// (c = E / ((1 + nu)(1 - 2 nu))), the node block (a, b) of Bq' De Bq is
// lambda g_a g_b' + mu g_b g_a' + mu (g_a . g_b) I for the gradients g of
// N_a, N_b at the IP, so
//     Ke_ab[i][j] = lambda S_ab[i][j] + mu S_ab[j][i] + mu tr(S_ab) delta_ij,
//     S_ab[i][j]  = sum_q coef_q g_a,i(q) g_b,j(q),
// one product G' (C G) whose inner dimension is the IPs, four per MFMA
// k-step (27 IPs: 7 k-steps).  Row (a, i) of S sits at 12 (a/4) + 4 i + a%4
// and column (b, j) at 16 j + b%16 (nodes 16..19 in a second pass), so
// that lane l holds all nine entries of S_ab for the nodes a = l/16 (mod 4)
// and b = l%16 (mod 16) and forms its blocks of Ke in registers.  Same
// operands as the exact kernel up to this algebra; the sums run in another
// order, so Ke agrees with it to rounding.
template <int NN, int NQ>
__device__ __attribute__((always_inline)) inline void
ke_gram_kernel_at(int64_t bid_, int64_t ne, const double *__restrict__ X, const int32_t *__restrict__ conn,
                  const int32_t *__restrict__ emat, const AmrMat *__restrict__ mats,
                  const double *__restrict__ dNdR, const double *__restrict__ wip, double *__restrict__ Ke,
                  int32_t *__restrict__ err)
{
    constexpr int KS = (NQ + 3) / 4;                   // k-steps over the IPs
    constexpr int NAG = (NN + 3) / 4;                  // row node groups: a = 4 ag + l/16
    constexpr int MT = (3 * (NAG - 1) + 2) / 4 + 1;    // row tiles
    constexpr int NP = (NN + 15) / 16;                 // column passes of 16 nodes
    const int64_t e = bid_;
    const int lane = threadIdx.x, lr = lane % 16, lk = lane / 16;
    __shared__ double sC[NN * 3];
    __shared__ double sInv[NQ][9];
    __shared__ double sCoef[4 * KS];
    __shared__ double sG[4 * KS][NN * 3];
    __shared__ int sBad;
    for (int i = lane; i < NN * 3; i += 64) sC[i] = X[(int64_t)conn[e * NN + i / 3] * 3 + i % 3];
    if (lane == 0) sBad = 0;
    __syncthreads();
    if (lane < NQ) {
        const double *dq = dNdR + lane * NN * 3;
        double J[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, inv[9];
#pragma unroll 2
        for (int n = 0; n < NN; n++) {
            const double d0 = dq[n * 3], d1 = dq[n * 3 + 1], d2 = dq[n * 3 + 2];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const double x = sC[n * 3 + i];
                J[i * 3] = fma(x, d0, J[i * 3]);
                J[i * 3 + 1] = fma(x, d1, J[i * 3 + 1]);
                J[i * 3 + 2] = fma(x, d2, J[i * 3 + 2]);
            }
        }
        const double det = inv3(J, inv);
        if (!(det > 0.0)) sBad = 1;
        for (int k = 0; k < 9; k++) sInv[lane][k] = inv[k];
        sCoef[lane] = det * wip[lane];
    } else if (lane < 4 * KS) {
        sCoef[lane] = 0.0;
    }
    __syncthreads();
    for (int t = lane; t < 4 * KS * NN; t += 64) {
        const int q = t / NN, n = t % NN;
        for (int j = 0; j < 3; j++) {
            double a = 0.0;
            if (q < NQ) {
                const double *dq = dNdR + (q * NN + n) * 3;
                for (int i = 0; i < 3; i++) a = fma(dq[i], sInv[q][i * 3 + j], a);
            }
            sG[q][n * 3 + j] = a;
        }
    }
    __syncthreads();
    // A operand rows: product row p = 16 mt + l%16 is (a, i) or padding
    int aoff[MT];
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
        const int p = 16 * mt + lr, u = p / 4, a = 4 * (u / 3) + p % 4;
        aoff[mt] = (u < 3 * NAG && a < NN) ? a * 3 + u % 3 : -1;
    }
    const AmrMat mat = mats[emat[e]];
    const double E = mat.p[0], nu = mat.p[1];
    const double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double lam = c * nu, mu = 0.5 * (c * (1.0 - 2.0 * nu));
    double *Ke_e = Ke + e * (int64_t)(NN * NN * 9);
#pragma unroll 1
    for (int pp = 0; pp < NP; pp++) {
        const int b = 16 * pp + lr;
        f64x4m acc[MT][3];
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int j = 0; j < 3; j++) acc[mt][j] = f64x4m{0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
        for (int s = 0; s < KS; s++) {
            const int q = 4 * s + lk;
            const double cq = sCoef[q];
            double av[MT], bv[3];
#pragma unroll
            for (int mt = 0; mt < MT; mt++) av[mt] = aoff[mt] >= 0 ? sG[q][aoff[mt]] : 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) bv[j] = b < NN ? cq * sG[q][b * 3 + j] : 0.0;
#pragma unroll
            for (int mt = 0; mt < MT; mt++)
#pragma unroll
                for (int j = 0; j < 3; j++)
                    acc[mt][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mt], bv[j], acc[mt][j], 0, 0, 0);
        }
        if (b < NN) {
#pragma unroll
            for (int ag = 0; ag < NAG; ag++) {
                const int a = 4 * ag + lk;
                if (a >= NN) continue;
                double S[3][3];
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        const int u = 3 * ag + i;
                        S[i][j] = acc[u / 4][j][u % 4];
                    }
                const double tr = S[0][0] + S[1][1] + S[2][2];
                double *blk = Ke_e + ((int64_t)a * NN + b) * 9;
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++)
                        blk[i * 3 + j] = fma(lam, S[i][j], mu * S[j][i]) + (i == j ? mu * tr : 0.0);
            }
        }
    }
    if (lane == 0 && sBad) atomicMin(err, (int32_t)(e + 1));
}

template <int NN, int NQ>
__global__ void __launch_bounds__(64) ke_gram_kernel(int64_t ne, const double *__restrict__ X,
                                  const int32_t *__restrict__ conn, const int32_t *__restrict__ emat,
                                  const AmrMat *__restrict__ mats, const double *__restrict__ dNdR,
                                  const double *__restrict__ wip, double *__restrict__ Ke, int32_t *__restrict__ err)
{
    ke_gram_kernel_at<NN, NQ>(blockIdx.x, ne, X, conn, emat, mats, dNdR, wip, Ke, err);
}

// ----------------------------------------------------------- element mass
//
// One 64-lane workgroup per element.  Lanes q < NQ build J (the stiffness
// kernel's fma chains) and coef = (rho * detJ) * w; then lane pairs (a, b)
// accumulate m_ab = m_ab + coef_q * (N_a N_b) over the IPs in order.  N'N
// couples equal directions only, so the node block of Me is diag(m_ab) and
// one value per node pair is stored (Me[(e*NN + a)*NN + b]); the six zeros
// of each block fall under mount_M's eps filter.

template <int NN, int NQ>
__global__ void __launch_bounds__(64) mass_kernel(int64_t ne, const double *__restrict__ X,
                                                  const int32_t *__restrict__ conn,
                                                  const double *__restrict__ rho,
                                                  const double *__restrict__ dNdR,
                                                  const double *__restrict__ Nip,
                                                  const double *__restrict__ wip, double *__restrict__ Me,
                                                  int32_t *__restrict__ err)
{
    const int64_t e = blockIdx.x;
    const int lane = threadIdx.x;
    __shared__ double sC[NN * 3];
    __shared__ double sN[NQ * NN];
    __shared__ double sCoef[NQ];
    __shared__ int sBad;
    for (int i = lane; i < NN * 3; i += 64) sC[i] = X[(int64_t)conn[e * NN + i / 3] * 3 + i % 3];
    for (int i = lane; i < NQ * NN; i += 64) sN[i] = Nip[i];
    if (lane == 0) sBad = 0;
    __syncthreads();
    if (lane < NQ) {
        const double *dq = dNdR + lane * NN * 3;
        double J[9], inv[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double a = 0.0;
                for (int n = 0; n < NN; n++) a = fma(sC[n * 3 + i], dq[n * 3 + j], a);
                J[i * 3 + j] = a;
            }
        const double det = inv3(J, inv);
        if (!(det > 0.0)) sBad = 1;
        sCoef[lane] = rho[e] * det * wip[lane];
    }
    __syncthreads();
    for (int pr = lane; pr < NN * NN; pr += 64) {
        const int a = pr / NN, b = pr % NN;
        double mab = 0.0;
        for (int q = 0; q < NQ; q++) mab += sCoef[q] * (sN[q * NN + a] * sN[q * NN + b]);
        Me[(e * NN + a) * NN + b] = mab;
    }
    if (lane == 0 && sBad) atomicMin(err, (int32_t)(e + 1));
}

// One thread per block: the block's element values m_ab (|m| >= eps) summed
// in element order from 0.0, on the three diagonal positions; the others are
// 0 and not in the pattern.  The same sums and flags as the generic block
// assembly over diag(m_ab) blocks.
__global__ void __launch_bounds__(256) mass_assemble_kernel(int64_t nblk, const int64_t *__restrict__ cptr,
                                                            const int32_t *__restrict__ csrc,
                                                            const double *__restrict__ Me, double *__restrict__ Mv,
                                                            uint8_t *__restrict__ flags)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nblk) return;
    double acc = 0.0;
    bool any = false;
    for (int64_t p = cptr[k]; p < cptr[k + 1]; p++) {
        const double x = Me[csrc[p]];
        if (!(fabs(x) < DEPS)) {
            acc += x;
            any = true;
        }
    }
#pragma unroll
    for (int s = 0; s < 9; s++) {
        const bool d = s == 0 || s == 4 || s == 8;
        Mv[kv_at(k, s)] = d ? acc : 0.0;
        flags[kv_at(k, s)] = d && any;
    }
}

// --------------------------------------------------------------- assembly
//
// One thread per K block: sum the element blocks listed for it in element
// order, each scalar only if |v| >= eps (mount_K's filter), starting from
// 0.0 (0.0 + v1 == v1).  Writes FP64 and FP32 values and, when asked, the
// survivor flags that define the reference's scalar pattern.  Values are
// stored in the kv_at(k, s) chunked layout.
__device__ __attribute__((always_inline)) inline void
assemble_kernel_at(int64_t bid_, int64_t nblk, const int64_t *__restrict__ cptr, const int32_t *__restrict__ csrc,
                   const double *__restrict__ Ke, double *__restrict__ Kv, float *__restrict__ Kv32,
                   uint8_t *__restrict__ flags)
{
    int64_t k = (int64_t)bid_ * blockDim.x + threadIdx.x;
    if (k >= nblk) return;
    double acc[9];
    unsigned fl = 0;
#pragma unroll
    for (int s = 0; s < 9; s++) acc[s] = 0.0;
    for (int64_t p = cptr[k]; p < cptr[k + 1]; p++) {
        const double *v = Ke + (int64_t)csrc[p] * 9;
#pragma unroll
        for (int s = 0; s < 9; s++) {
            double x = v[s];
            if (!(fabs(x) < DEPS)) {
                acc[s] += x;
                fl |= 1u << s;
            }
        }
    }
#pragma unroll
    for (int s = 0; s < 9; s++) Kv[kv_at(k, s)] = acc[s];
    if (Kv32) {
#pragma unroll
        for (int s = 0; s < 9; s++) Kv32[kv32_at(k, s)] = (float)acc[s];
    }
    if (flags) {
#pragma unroll
        for (int s = 0; s < 9; s++) flags[kv_at(k, s)] = (fl >> s) & 1u;
    }
}

__global__ void __launch_bounds__(256) assemble_kernel(int64_t nblk, const int64_t *__restrict__ cptr,
                                  const int32_t *__restrict__ csrc, const double *__restrict__ Ke,
                                  double *__restrict__ Kv, float *__restrict__ Kv32, uint8_t *__restrict__ flags)
{
    assemble_kernel_at(blockIdx.x, nblk, cptr, csrc, Ke, Kv, Kv32, flags);
}

// Kp = a K + b M in K's storage (the pseudo-stiffness of the dynamic
// solver, src/mech/dyn-solver.jl:330-331): one thread per block; an entry is
// in the pattern when it is in K's or in M's; FP64 and FP32 values written.
__global__ void __launch_bounds__(256) combine_km_kernel(int64_t nblk, double a, double b, double *__restrict__ Kv,
                                                         float *__restrict__ Kv32, uint8_t *__restrict__ Kflag,
                                                         const double *__restrict__ Mv,
                                                         const uint8_t *__restrict__ Mflag)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= nblk) return;
#pragma unroll
    for (int s = 0; s < 9; s++) {
        const int64_t i = kv_at(k, s);
        const double v = a * Kv[i] + b * Mv[i];
        Kv[i] = v;
        Kv32[kv32_at(k, s)] = (float)v;
        Kflag[i] = Kflag[i] | Mflag[i];
    }
}

// K := a1 K1 + a2 K (the predictor-corrector tangent of mech_stage_solver!,
// src/mech/mech-solver.jl:343: K = a1*K + a2*K2, K1 the predictor's tangent
// saved before K2 was assembled over it): one thread per block; FP64 and
// FP32 values written.
__global__ void __launch_bounds__(256) combine_kk_kernel(int64_t nblk, double a1, const double *__restrict__ K1,
                                                         double a2, double *__restrict__ Kv,
                                                         float *__restrict__ Kv32)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= nblk) return;
#pragma unroll
    for (int s = 0; s < 9; s++) {
        const int64_t i = kv_at(k, s);
        const double v = a1 * K1[i] + a2 * Kv[i];
        Kv[i] = v;
        Kv32[kv32_at(k, s)] = (float)v;
    }
}

// Jacobi inverse diagonal (0 at prescribed dofs)
__device__ __attribute__((always_inline)) inline void
jacobi_kernel_at(int64_t bid_, int64_t nnodes, int64_t nblk, const int32_t *__restrict__ diag,
                 const double *__restrict__ Kv, const double *__restrict__ mask, double *__restrict__ dinv,
                 float *__restrict__ dinv32)
{
    int64_t i = (int64_t)bid_ * blockDim.x + threadIdx.x;
    if (i >= nnodes * 3) return;
    int64_t n = i / 3;
    int d = (int)(i % 3);
    const int32_t db = diag[n];  // -1 for a node in no element (it has no dofs)
    double kd = db >= 0 ? Kv[kv_at(db, d * 4)] : 1.0;
    bool fr = mask[i] != 0.0 && db >= 0;
    dinv[i] = fr ? 1.0 / kd : 0.0;
    dinv32[i] = fr ? 1.0f / (float)kd : 0.0f;
}

__global__ void jacobi_kernel(int64_t nnodes, int64_t nblk, const int32_t *__restrict__ diag,
                              const double *__restrict__ Kv, const double *__restrict__ mask,
                              double *__restrict__ dinv, float *__restrict__ dinv32)
{
    jacobi_kernel_at(blockIdx.x, nnodes, nblk, diag, Kv, mask, dinv, dinv32);
}

// 3x3 nodal block Jacobi: inverse of each node's (symmetrised) diagonal
// block restricted to the node's free dofs; rows/columns of prescribed dofs
// are zero.
__device__ __attribute__((always_inline)) inline void
block_jacobi_kernel_at(int64_t bid_, int64_t nnodes, int64_t nblk, const int32_t *__restrict__ diag,
                       const double *__restrict__ Kv, const double *__restrict__ mask, double *__restrict__ binv,
                       float *__restrict__ binv32)
{
    const int64_t n = (int64_t)bid_ * blockDim.x + threadIdx.x;
    if (n >= nnodes) return;
    const int32_t db = diag[n];
    double A[3][3], I[3][3];
    int fr[3], nf = 0, idx[3];
    for (int r = 0; r < 3; r++) {
        fr[r] = db >= 0 && mask[n * 3 + r] != 0.0;
        if (fr[r]) idx[nf++] = r;
        for (int c = 0; c < 3; c++) {
            A[r][c] = db >= 0 ? 0.5 * (Kv[kv_at(db, r * 3 + c)] + Kv[kv_at(db, c * 3 + r)]) : 0.0;
            I[r][c] = 0.0;
        }
    }
    if (nf == 1) {
        const int a = idx[0];
        I[a][a] = 1.0 / A[a][a];
    } else if (nf == 2) {
        const int a = idx[0], b = idx[1];
        const double det = A[a][a] * A[b][b] - A[a][b] * A[b][a];
        I[a][a] = A[b][b] / det;
        I[b][b] = A[a][a] / det;
        I[a][b] = -A[a][b] / det;
        I[b][a] = -A[b][a] / det;
    } else if (nf == 3) {
        double J[9], inv[9];
        for (int k = 0; k < 9; k++) J[k] = A[k / 3][k % 3];
        inv3(J, inv);
        for (int k = 0; k < 9; k++) I[k / 3][k % 3] = inv[k];
    }
    for (int k = 0; k < 9; k++) {
        binv[n * 9 + k] = I[k / 3][k % 3];
        binv32[n * 9 + k] = (float)I[k / 3][k % 3];
    }
}

__global__ void block_jacobi_kernel(int64_t nnodes, int64_t nblk, const int32_t *__restrict__ diag,
                                    const double *__restrict__ Kv, const double *__restrict__ mask,
                                    double *__restrict__ binv, float *__restrict__ binv32)
{
    block_jacobi_kernel_at(blockIdx.x, nnodes, nblk, diag, Kv, mask, binv, binv32);
}

// ------------------------------------------------------------- IP update
//
// EPW elements per 256-thread workgroup.  Phase A: one thread per IP builds
// J, its inverse and coef, streams the strain increment deps = B dU over the
// element's nodes (the oracle's column-ascending fma chains), runs the
// constitutive update from the committed state into the current state and
// keeps (inv J, coef, dsig) in LDS.  Phase B: one thread per element column
// accumulates dF[c] = fma(coef_q, (B' dsig_q)[c], dF[c]) over the IPs in
// order.
template <int NN, int NQ, int EPW>
__device__ __attribute__((always_inline)) inline void
update_kernel_at(int64_t bid_, int64_t ne, const double *__restrict__ X, const int32_t *__restrict__ conn,
                 const int32_t *__restrict__ emat, const AmrMat *__restrict__ mats, const double *__restrict__ dNdR,
                 const double *__restrict__ wip, const double *__restrict__ st_in, double *__restrict__ st_out,
                 int64_t nip, const double *__restrict__ dU, double *__restrict__ dF, int32_t *__restrict__ estat,
                 int32_t *__restrict__ err)
{
    constexpr int ND = 3 * NN;
    __shared__ double sC[EPW][NN * 3];
    __shared__ double sU[EPW][ND];
    __shared__ double sInv[EPW * NQ][9];
    __shared__ double sCoef[EPW * NQ];
    __shared__ double sDs[EPW * NQ][6];
    __shared__ int sStat[EPW * NQ];
    const int64_t e0 = (int64_t)bid_ * EPW;
    const int tid = threadIdx.x;

    for (int i = tid; i < EPW * NN * 3; i += blockDim.x) {
        int el = i / (NN * 3), r = i % (NN * 3);
        int64_t e = e0 + el;
        if (e < ne) {
            int64_t node = conn[e * NN + r / 3];
            sC[el][r] = X[node * 3 + r % 3];
            sU[el][r] = dU[node * 3 + r % 3];
        }
    }
    __syncthreads();

    if (tid < EPW * NQ) {
        const int el = tid / NQ, q = tid % NQ;
        const int64_t e = e0 + el;
        int stat = 0;
        if (e < ne) {
            const double *dq = dNdR + q * NN * 3;
            double J[9], inv[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double a = 0.0;
                    for (int n = 0; n < NN; n++) a = fma(sC[el][n * 3 + i], dq[n * 3 + j], a);
                    J[i * 3 + j] = a;
                }
            double det = inv3(J, inv);
            if (!(det > 0.0)) stat = 2;
            double de[6] = {0, 0, 0, 0, 0, 0};
            for (int n = 0; n < NN; n++) {
                double g[3];
                for (int j = 0; j < 3; j++) {
                    double a = 0.0;
                    for (int i = 0; i < 3; i++) a = fma(dq[n * 3 + i], inv[i * 3 + j], a);
                    g[j] = a;
                }
                const double hx = g[0] / SR2, hy = g[1] / SR2, hz = g[2] / SR2;
                const double u0 = sU[el][n * 3], u1 = sU[el][n * 3 + 1], u2 = sU[el][n * 3 + 2];
                de[0] = fma(g[0], u0, de[0]);
                de[1] = fma(g[1], u1, de[1]);
                de[2] = fma(g[2], u2, de[2]);
                de[3] = fma(hy, u2, fma(hz, u1, de[3]));
                de[4] = fma(hx, u2, fma(hz, u0, de[4]));
                de[5] = fma(hx, u1, fma(hy, u0, de[5]));
            }
            const int64_t ip = e * NQ + q;
            double s14[AMR_NSTATE], ds[6];
            for (int f = 0; f < AMR_NSTATE; f++) s14[f] = st_in[f * nip + ip];
            const AmrMat m = mats[emat[e]];
            if (update_ip(m, s14, de, ds)) stat = stat ? stat : 1;
            for (int f = 0; f < AMR_NSTATE; f++) st_out[f * nip + ip] = s14[f];
            for (int i = 0; i < 9; i++) sInv[tid][i] = inv[i];
            sCoef[tid] = det * wip[q];
            for (int k = 0; k < 6; k++) sDs[tid][k] = ds[k];
        }
        sStat[tid] = stat;
    }
    __syncthreads();

    for (int it = tid; it < EPW * ND; it += blockDim.x) {
        const int el = it / ND, c = it % ND;
        const int64_t e = e0 + el;
        if (e >= ne) continue;
        const int n = c / 3, d = c % 3;
        double F = 0.0;
        for (int q = 0; q < NQ; q++) {
            const int slot = el * NQ + q;
            const double *dq = dNdR + q * NN * 3;
            double g[3];
            for (int j = 0; j < 3; j++) {
                double a = 0.0;
                for (int i = 0; i < 3; i++) a = fma(dq[n * 3 + i], sInv[slot][i * 3 + j], a);
                g[j] = a;
            }
            const double *s = sDs[slot];
            double t;
            if (d == 0) t = fma(g[1] / SR2, s[5], fma(g[2] / SR2, s[4], fma(g[0], s[0], 0.0)));
            else if (d == 1) t = fma(g[0] / SR2, s[5], fma(g[2] / SR2, s[3], fma(g[1], s[1], 0.0)));
            else t = fma(g[0] / SR2, s[4], fma(g[1] / SR2, s[3], fma(g[2], s[2], 0.0)));
            F = fma(sCoef[slot], t, F);
        }
        dF[e * ND + c] = F;
    }
    if (tid < EPW) {
        const int64_t e = e0 + tid;
        if (e < ne) {
            int s = 0;
            for (int q = 0; q < NQ; q++) {
                int v = sStat[tid * NQ + q];
                if (v) {
                    s = v;
                    break;
                }
            }
            estat[e] = s;
            if (s == 1) atomicMin(err + 1, (int32_t)(e + 1));
            if (s == 2) atomicMin(err + 2, (int32_t)(e + 1));
        }
    }
}

template <int NN, int NQ, int EPW>
__global__ void __launch_bounds__(256) update_kernel(int64_t ne, const double *__restrict__ X,
                                  const int32_t *__restrict__ conn, const int32_t *__restrict__ emat,
                                  const AmrMat *__restrict__ mats, const double *__restrict__ dNdR,
                                  const double *__restrict__ wip, const double *__restrict__ st_in,
                                  double *__restrict__ st_out, int64_t nip, const double *__restrict__ dU,
                                  double *__restrict__ dF, int32_t *__restrict__ estat, int32_t *__restrict__ err)
{
    update_kernel_at<NN, NQ, EPW>(blockIdx.x, ne, X, conn, emat, mats, dNdR, wip, st_in, st_out, nip, dU, dF, estat,
                                  err);
}

// dFin[3n+d] = sum over (e, la) incident to node n, in element order
__device__ __attribute__((always_inline)) inline void
force_gather_kernel_at(int64_t bid_, int64_t nnodes, int nn, const int32_t *__restrict__ fptr,
                       const int32_t *__restrict__ fsrc, const double *__restrict__ dF, double *__restrict__ out)
{
    int64_t i = (int64_t)bid_ * blockDim.x + threadIdx.x;
    if (i >= nnodes * 3) return;
    int64_t n = i / 3;
    int d = (int)(i % 3);
    double acc = 0.0;
    for (int32_t p = fptr[n]; p < fptr[n + 1]; p++) acc += dF[(int64_t)fsrc[p] * 3 + d];
    out[i] = acc;
}

__global__ void force_gather_kernel(int64_t nnodes, int nn, const int32_t *__restrict__ fptr,
                                    const int32_t *__restrict__ fsrc, const double *__restrict__ dF,
                                    double *__restrict__ out)
{
    force_gather_kernel_at(blockIdx.x, nnodes, nn, fptr, fsrc, dF, out);
}

template <int NN, int NQ, int ACH>
int launch_ke_t(amr_model *m)
{
    if (m->ne == 0) return AMR_OK;
    hipLaunchKernelGGL((ke_kernel<NN, NQ, ACH>), dim3((unsigned)m->ne), dim3(64), 0, m->stream, m->ne, m->X,
                       m->conn, m->emat, m->mats, m->dNdR, m->wip, m->st_cur, m->nip_total,
                       m->all_elastic ? 1 : 0, m->Ke, m->err);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

template <int NN, int NQ, int EPW>
int launch_update_t(amr_model *m, const double *dU)
{
    if (m->ne == 0) return AMR_OK;
    unsigned grid = (unsigned)((m->ne + EPW - 1) / EPW);
    hipLaunchKernelGGL((update_kernel<NN, NQ, EPW>), dim3(grid), dim3(256), 0, m->stream, m->ne, m->X, m->conn,
                       m->emat, m->mats, m->dNdR, m->wip, m->st_com, m->st_cur, m->nip_total, dU, m->dF,
                       m->estat, m->err);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

// Element tiles.  Per element shape (HEX8 8 IP, HEX20 27 IP, HEX20 8 IP,
// TET10 4 IP) three stiffness tilings (ACH row nodes per lane; NN * ceil(NN
// / ACH) <= 64 lanes) and three update tilings (EPW elements per 256-thread
// workgroup, EPW * NQ <= 256).  Each Ke entry and each IP update is the
// same arithmetic in every tiling, so the choice changes speed only.
constexpr int ACH_OPT[4][3] = {{1, 2, 4}, {7, 10, 20}, {7, 10, 20}, {2, 5, 10}};
constexpr int EPW_OPT[4][3] = {{32, 16, 8}, {9, 4, 2}, {16, 8, 32}, {32, 16, 64}};

int shape_index(const amr_model *m)
{
    switch (m->etype) {
    case AMR_HEX8: return 0;
    case AMR_HEX20: return m->nips == 27 ? 1 : 2;
    case AMR_TET10: return 3;
    }
    return -1;
}

// the process-wide choice of the tuner per shape (-1: not tuned)
std::mutex g_tile_mu;
int g_ke_var[4] = {-1, -1, -1, -1}, g_up_var[4] = {-1, -1, -1, -1};

int ke_variant(const amr_model *m, int sh)
{
    if (m->ke_var >= 0) return m->ke_var;
    std::lock_guard<std::mutex> lk(g_tile_mu);
    return g_ke_var[sh] >= 0 ? g_ke_var[sh] : 0;
}

int up_variant(const amr_model *m, int sh)
{
    if (m->up_var >= 0) return m->up_var;
    std::lock_guard<std::mutex> lk(g_tile_mu);
    return g_up_var[sh] >= 0 ? g_up_var[sh] : 0;
}

int launch_ke_var(amr_model *m, int sh, int v)
{
#define KE3(NN, NQ, A0, A1, A2) \
    return v == 0 ? launch_ke_t<NN, NQ, A0>(m) : v == 1 ? launch_ke_t<NN, NQ, A1>(m) : launch_ke_t<NN, NQ, A2>(m)
    switch (sh) {
    case 0: KE3(8, 8, 1, 2, 4);
    case 1: KE3(20, 27, 7, 10, 20);
    case 2: KE3(20, 8, 7, 10, 20);
    case 3: KE3(10, 4, 2, 5, 10);
    }
#undef KE3
    amr_set_error("unknown element type");
    return AMR_ERR_ARG;
}

int launch_update_var(amr_model *m, int sh, int v, const double *dU)
{
#define UP3(NN, NQ, E0, E1, E2)                                                                        \
    return v == 0 ? launch_update_t<NN, NQ, E0>(m, dU)                                               \
                  : v == 1 ? launch_update_t<NN, NQ, E1>(m, dU) : launch_update_t<NN, NQ, E2>(m, dU)
    switch (sh) {
    case 0: UP3(8, 8, 32, 16, 8);
    case 1: UP3(20, 27, 9, 4, 2);
    case 2: UP3(20, 8, 16, 8, 32);
    case 3: UP3(10, 4, 32, 16, 64);
    }
#undef UP3
    amr_set_error("unknown element type");
    return AMR_ERR_ARG;
}

}  // namespace

int amr_launch_ke_path(amr_model *m, int path)
{
    const int sh = shape_index(m);
    if (sh < 0) {
        amr_set_error("unknown element type");
        return AMR_ERR_ARG;
    }
    if (path == AMR_KE_MFMA && m->all_elastic) {
        if (m->ne == 0) return AMR_OK;
        const dim3 g((unsigned)m->ne), b(64);
        if (m->etype == AMR_HEX20 && m->nips == 27)
            hipLaunchKernelGGL((ke_gram_kernel<20, 27>), g, b, 0, m->stream, m->ne, m->X, m->conn, m->emat, m->mats,
                               m->dNdR, m->wip, m->Ke, m->err);
        else if (m->etype == AMR_HEX20 && m->nips == 8)
            hipLaunchKernelGGL((ke_gram_kernel<20, 8>), g, b, 0, m->stream, m->ne, m->X, m->conn, m->emat, m->mats,
                               m->dNdR, m->wip, m->Ke, m->err);
        else if (m->etype == AMR_HEX8 && m->nips == 8)
            hipLaunchKernelGGL((ke_gram_kernel<8, 8>), g, b, 0, m->stream, m->ne, m->X, m->conn, m->emat, m->mats,
                               m->dNdR, m->wip, m->Ke, m->err);
        else if (m->etype == AMR_TET10 && m->nips == 4)
            hipLaunchKernelGGL((ke_gram_kernel<10, 4>), g, b, 0, m->stream, m->ne, m->X, m->conn, m->emat, m->mats,
                               m->dNdR, m->wip, m->Ke, m->err);
        else
            return launch_ke_var(m, sh, ke_variant(m, sh));
        AMR_HIP(hipGetLastError());
        return AMR_OK;
    }
    if (path == AMR_KE_MFMA && m->etype == AMR_HEX20) {
        if (m->ne == 0) return AMR_OK;
        if (m->nips == 27)
            hipLaunchKernelGGL((ke_mfma_hex20_kernel<27, 9>), dim3((unsigned)m->ne), dim3(64), 0, m->stream, m->ne, m->X,
                               m->conn, m->emat, m->mats, m->dNdR, m->wip, m->st_cur, m->nip_total,
                               m->all_elastic ? 1 : 0, m->Ke, m->err);
        else
            hipLaunchKernelGGL((ke_mfma_hex20_kernel<8, 8>), dim3((unsigned)m->ne), dim3(64), 0, m->stream, m->ne, m->X,
                               m->conn, m->emat, m->mats, m->dNdR, m->wip, m->st_cur, m->nip_total,
                               m->all_elastic ? 1 : 0, m->Ke, m->err);
        AMR_HIP(hipGetLastError());
        return AMR_OK;
    }
    return launch_ke_var(m, sh, ke_variant(m, sh));
}

int amr_launch_ke(amr_model *m) { return amr_launch_ke_path(m, m->ke_path); }

int amr_elem_tiles(amr_model *m, int *ach, int *epw)
{
    const int sh = shape_index(m);
    if (sh < 0) return AMR_ERR_ARG;
    *ach = ACH_OPT[sh][ke_variant(m, sh)];
    *epw = EPW_OPT[sh][up_variant(m, sh)];
    return AMR_OK;
}

int amr_elem_tiles_set(amr_model *m, int ach, int epw)
{
    const int sh = shape_index(m);
    if (sh < 0) return AMR_ERR_ARG;
    int kv = -1, uv = -1;
    for (int i = 0; i < 3; i++) {
        if (ACH_OPT[sh][i] == ach) kv = i;
        if (EPW_OPT[sh][i] == epw) uv = i;
    }
    if ((ach != 0 && kv < 0) || (epw != 0 && uv < 0)) {
        amr_set_error("amr_set_elem_tiles: no such tiling for this element shape");
        return AMR_ERR_ARG;
    }
    m->ke_var = ach == 0 ? -1 : kv;
    m->up_var = epw == 0 ? -1 : uv;
    return AMR_OK;
}

// Times the three tilings of each kernel on the model (element stiffness
// from the current states; the update from the committed states with a zero
// displacement increment, written to the current states) and keeps the
// fastest for this model and, process-wide, for its element shape.
int amr_elem_tiles_tune(amr_model *m, int *ach, int *epw)
{
    const int sh = shape_index(m);
    if (sh < 0) return AMR_ERR_ARG;
    hipEvent_t e0, e1;
    AMR_HIP(hipEventCreate(&e0));
    AMR_HIP(hipEventCreate(&e1));
    struct Ev {
        hipEvent_t a, b;
        ~Ev()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } guard{e0, e1};
    double *dU = m->v[6];
    const size_t n3 = (size_t)m->nnodes * 3;
    AMR_HIP(hipMemsetAsync(dU, 0, n3 * sizeof(double), m->stream));
    AMR_HIP(hipMemsetAsync(m->err, 0x7f, 3 * sizeof(int32_t), m->stream));
    auto best_of = [&](auto &&launch, int &best) -> int {
        float tbest = 0.0f;
        best = 0;
        for (int v = 0; v < 3; v++) {
            if (int rc = launch(v)) return rc;   // warm-up
            AMR_HIP(hipEventRecord(e0, m->stream));
            for (int r = 0; r < 3; r++)
                if (int rc = launch(v)) return rc;
            AMR_HIP(hipEventRecord(e1, m->stream));
            AMR_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            AMR_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (v == 0 || ms < tbest) {
                tbest = ms;
                best = v;
            }
        }
        return AMR_OK;
    };
    int kb = 0, ub = 0;
    if (int rc = best_of([&](int v) { return launch_ke_var(m, sh, v); }, kb)) return rc;
    if (int rc = best_of([&](int v) { return launch_update_var(m, sh, v, dU); }, ub)) return rc;
    AMR_HIP(hipMemsetAsync(m->err, 0x7f, 3 * sizeof(int32_t), m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    m->ke_var = kb;
    m->up_var = ub;
    {
        std::lock_guard<std::mutex> lk(g_tile_mu);
        g_ke_var[sh] = kb;
        g_up_var[sh] = ub;
    }
    *ach = ACH_OPT[sh][kb];
    *epw = EPW_OPT[sh][ub];
    return AMR_OK;
}

int amr_launch_assemble(amr_model *m, bool flags)
{
    m->k_version++;
    if (m->pat.nblk == 0) return AMR_OK;
    unsigned grid = (unsigned)((m->pat.nblk + 255) / 256);
    hipLaunchKernelGGL(assemble_kernel, dim3(grid), dim3(256), 0, m->stream, m->pat.nblk, m->cptr, m->csrc, m->Ke,
                       m->Kv, m->Kv32, flags ? m->Kflag : nullptr);
    AMR_HIP(hipGetLastError());
    unsigned g2 = (unsigned)((m->nnodes * 3 + 255) / 256);
    hipLaunchKernelGGL(jacobi_kernel, dim3(g2), dim3(256), 0, m->stream, m->nnodes, m->pat.nblk, m->diag, m->Kv, m->mask,
                       m->dinv, m->dinv32);
    AMR_HIP(hipGetLastError());
    const unsigned g3 = (unsigned)((m->nnodes + 255) / 256);
    hipLaunchKernelGGL(block_jacobi_kernel, dim3(g3), dim3(256), 0, m->stream, m->nnodes, m->pat.nblk, m->diag, m->Kv,
                       m->mask, m->binv, m->binv32);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

int amr_launch_combine_km(amr_model *m, double a, double b)
{
    m->k_version++;
    if (m->pat.nblk == 0) return AMR_OK;
    const unsigned grid = (unsigned)((m->pat.nblk + 255) / 256);
    hipLaunchKernelGGL(combine_km_kernel, dim3(grid), dim3(256), 0, m->stream, m->pat.nblk, a, b, m->Kv, m->Kv32,
                       m->Kflag, m->Mv, m->Mflag);
    AMR_HIP(hipGetLastError());
    const unsigned g2 = (unsigned)((m->nnodes * 3 + 255) / 256);
    hipLaunchKernelGGL(jacobi_kernel, dim3(g2), dim3(256), 0, m->stream, m->nnodes, m->pat.nblk, m->diag, m->Kv, m->mask,
                       m->dinv, m->dinv32);
    AMR_HIP(hipGetLastError());
    const unsigned g3 = (unsigned)((m->nnodes + 255) / 256);
    hipLaunchKernelGGL(block_jacobi_kernel, dim3(g3), dim3(256), 0, m->stream, m->nnodes, m->pat.nblk, m->diag, m->Kv,
                       m->mask, m->binv, m->binv32);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

int amr_launch_combine_kk(amr_model *m, double a1, const double *K1, double a2)
{
    m->k_version++;
    if (m->pat.nblk == 0) return AMR_OK;
    const unsigned grid = (unsigned)((m->pat.nblk + 255) / 256);
    hipLaunchKernelGGL(combine_kk_kernel, dim3(grid), dim3(256), 0, m->stream, m->pat.nblk, a1, K1, a2, m->Kv,
                       m->Kv32);
    AMR_HIP(hipGetLastError());
    const unsigned g2 = (unsigned)((m->nnodes * 3 + 255) / 256);
    hipLaunchKernelGGL(jacobi_kernel, dim3(g2), dim3(256), 0, m->stream, m->nnodes, m->pat.nblk, m->diag, m->Kv, m->mask,
                       m->dinv, m->dinv32);
    AMR_HIP(hipGetLastError());
    const unsigned g3 = (unsigned)((m->nnodes + 255) / 256);
    hipLaunchKernelGGL(block_jacobi_kernel, dim3(g3), dim3(256), 0, m->stream, m->nnodes, m->pat.nblk, m->diag, m->Kv,
                       m->mask, m->binv, m->binv32);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

template <int NN, int NQ>
int launch_mass_t(amr_model *m)
{
    hipLaunchKernelGGL((mass_kernel<NN, NQ>), dim3((unsigned)m->ne), dim3(64), 0, m->stream, m->ne, m->X, m->conn,
                       m->rho_d, m->dNdR, m->Nd, m->wip, m->Me, m->err);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

int amr_launch_mass(amr_model *m)
{
    if (m->ne > 0) {
        int rc;
        switch (m->etype) {
        case AMR_HEX8: rc = launch_mass_t<8, 8>(m); break;
        case AMR_HEX20: rc = m->nips == 27 ? launch_mass_t<20, 27>(m) : launch_mass_t<20, 8>(m); break;
        case AMR_TET10: rc = launch_mass_t<10, 4>(m); break;
        default: amr_set_error("unknown element type"); return AMR_ERR_ARG;
        }
        if (rc) return rc;
    }
    if (m->pat.nblk == 0) return AMR_OK;
    const unsigned grid = (unsigned)((m->pat.nblk + 255) / 256);
    hipLaunchKernelGGL(mass_assemble_kernel, dim3(grid), dim3(256), 0, m->stream, m->pat.nblk, m->cptr, m->csrc,
                       m->Me, m->Mv, m->Mflag);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

int amr_launch_update(amr_model *m, const double *dU, double *dFin)
{
    const int sh = shape_index(m);
    if (sh < 0) {
        amr_set_error("unknown element type");
        return AMR_ERR_ARG;
    }
    if (int rc = launch_update_var(m, sh, up_variant(m, sh), dU)) return rc;
    unsigned g = (unsigned)((m->nnodes * 3 + 255) / 256);
    hipLaunchKernelGGL(force_gather_kernel, dim3(g), dim3(256), 0, m->stream, m->nnodes, m->nn, m->fptr, m->fsrc,
                       m->dF, dFin);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

// ------------------------------------------------ several models per launch
//
// The element phase of a set of models (a pool's members starting or ending
// a Newton-Raphson iteration together): each launch covers every model's
// workgroups, in model order; a workgroup finds its model in the launch's
// start table (binary search) and runs the single-model kernel's body on
// its own workgroup index, so every value is computed as in the model's own
// launch.

namespace {

#define AMR_AS1 __attribute__((address_space(1)))
// Pointer fields in the global address space: a kernel converts them to
// the bodies' generic parameters with one cast the compiler sees through,
// so the bodies use global (not flat) memory instructions, as with kernel
// arguments.
struct ElemJob {
    int64_t ne, nip, nblk, nnodes;
    int nn, all_elastic;
    const AMR_AS1 double *X;
    const AMR_AS1 int32_t *conn, *emat;
    const AMR_AS1 AmrMat *mats;
    const AMR_AS1 double *dNdR, *wip, *st_cur, *st_com;
    AMR_AS1 double *st_out, *Ke;
    AMR_AS1 int32_t *err;
    const AMR_AS1 int64_t *cptr;
    const AMR_AS1 int32_t *csrc, *diag, *fptr, *fsrc;
    AMR_AS1 double *Kv;
    AMR_AS1 float *Kv32;
    const AMR_AS1 double *mask;
    AMR_AS1 double *dinv, *binv;
    AMR_AS1 float *dinv32, *binv32;
    const AMR_AS1 double *dU;
    AMR_AS1 double *dF, *dFin;
    AMR_AS1 int32_t *estat;
};

// The launch's start table holds n + 1 workgroup offsets, then (from entry
// n + 1 on) the job of every 1,024th workgroup: a workgroup reads that hint
// and steps forward over the (few) jobs that start before it, a short chain
// of dependent loads where a binary search would take log2(n).
constexpr int HINT_SHIFT = 10;

__device__ inline int job_of(const int64_t *__restrict__ start, int n, int64_t b)
{
    int j = (int)start[n + 1 + (b >> HINT_SHIFT)];
    while (j + 1 < n && start[j + 1] <= b) j++;
    return j;
}

template <typename T>
__device__ inline T *gen(AMR_AS1 T *p)
{
    return (T *)p;
}

#define AMR_JOB(J, start, n)                                  \
    const int j_ = job_of(start, n, (int64_t)blockIdx.x);     \
    const ElemJob &q = J[j_];                                 \
    const int64_t b_ = (int64_t)blockIdx.x - start[j_]

// the error words of every model: first failing element "none" (the bytes
// 0x7f of the single-model reset)
__global__ void err_reset_kernel(const ElemJob *__restrict__ J, int n)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n)
        for (int k = 0; k < 3; k++) J[j].err[k] = 0x7f7f7f7f;
}

template <int NN, int NQ, int ACH>
__global__ void __launch_bounds__(64) ke_batch_kernel(const ElemJob *__restrict__ J, const int64_t *__restrict__ start,
                                                      int n)
{
    AMR_JOB(J, start, n);
    ke_kernel_at<NN, NQ, ACH>(b_, q.ne, gen(q.X), gen(q.conn), gen(q.emat), gen(q.mats), gen(q.dNdR), gen(q.wip),
                              gen(q.st_cur), q.nip, q.all_elastic,
                              gen(q.Ke), gen(q.err));
}

template <int NN, int NQ>
__global__ void __launch_bounds__(64) ke_gram_batch_kernel(const ElemJob *__restrict__ J,
                                                           const int64_t *__restrict__ start, int n)
{
    AMR_JOB(J, start, n);
    ke_gram_kernel_at<NN, NQ>(b_, q.ne, gen(q.X), gen(q.conn), gen(q.emat), gen(q.mats), gen(q.dNdR), gen(q.wip),
                              gen(q.Ke), gen(q.err));
}

template <int NQ, int NQC>
__global__ void __launch_bounds__(64, 2) ke_mfma_hex20_batch_kernel(const ElemJob *__restrict__ J,
                                                                    const int64_t *__restrict__ start, int n)
{
    AMR_JOB(J, start, n);
    ke_mfma_hex20_kernel_at<NQ, NQC>(b_, q.ne, gen(q.X), gen(q.conn), gen(q.emat), gen(q.mats), gen(q.dNdR),
                                     gen(q.wip), gen(q.st_cur), q.nip,
                                     q.all_elastic, gen(q.Ke), gen(q.err));
}

__global__ void __launch_bounds__(256) assemble_batch_kernel(const ElemJob *__restrict__ J,
                                                             const int64_t *__restrict__ start, int n)
{
    AMR_JOB(J, start, n);
    assemble_kernel_at(b_, q.nblk, gen(q.cptr), gen(q.csrc), gen(q.Ke), gen(q.Kv), gen(q.Kv32), nullptr);
}

__global__ void __launch_bounds__(256) jacobi_batch_kernel(const ElemJob *__restrict__ J,
                                                           const int64_t *__restrict__ start, int n)
{
    AMR_JOB(J, start, n);
    jacobi_kernel_at(b_, q.nnodes, q.nblk, gen(q.diag), gen(q.Kv), gen(q.mask), gen(q.dinv), gen(q.dinv32));
}

__global__ void __launch_bounds__(256) block_jacobi_batch_kernel(const ElemJob *__restrict__ J,
                                                                 const int64_t *__restrict__ start, int n)
{
    AMR_JOB(J, start, n);
    block_jacobi_kernel_at(b_, q.nnodes, q.nblk, gen(q.diag), gen(q.Kv), gen(q.mask), gen(q.binv), gen(q.binv32));
}

template <int NN, int NQ, int EPW>
__global__ void __launch_bounds__(256) update_batch_kernel(const ElemJob *__restrict__ J,
                                                           const int64_t *__restrict__ start, int n)
{
    AMR_JOB(J, start, n);
    update_kernel_at<NN, NQ, EPW>(b_, q.ne, gen(q.X), gen(q.conn), gen(q.emat), gen(q.mats), gen(q.dNdR), gen(q.wip),
                                  gen(q.st_com), gen(q.st_out), q.nip,
                                  gen(q.dU), gen(q.dF), gen(q.estat), gen(q.err));
}

__global__ void __launch_bounds__(256) force_gather_batch_kernel(const ElemJob *__restrict__ J,
                                                                 const int64_t *__restrict__ start, int n)
{
    AMR_JOB(J, start, n);
    force_gather_kernel_at(b_, q.nnodes, q.nn, gen(q.fptr), gen(q.fsrc), gen(q.dF), gen(q.dFin));
}

#undef AMR_JOB

template <typename T>
AMR_AS1 T *as1(T *p)
{
    return (AMR_AS1 T *)p;
}

ElemJob elem_job(const amr_model *m)
{
    ElemJob q{};
    q.ne = m->ne;
    q.nip = m->nip_total;
    q.nblk = m->pat.nblk;
    q.nnodes = m->nnodes;
    q.nn = m->nn;
    q.all_elastic = m->all_elastic ? 1 : 0;
    q.X = as1(m->X);
    q.conn = as1(m->conn);
    q.emat = as1(m->emat);
    q.mats = as1(m->mats);
    q.dNdR = as1(m->dNdR);
    q.wip = as1(m->wip);
    q.st_cur = as1(m->st_cur);
    q.st_com = as1(m->st_com);
    q.st_out = as1(m->st_cur);
    q.Ke = as1(m->Ke);
    q.err = as1(m->err);
    q.cptr = as1(m->cptr);
    q.csrc = as1(m->csrc);
    q.diag = as1(m->diag);
    q.fptr = as1(m->fptr);
    q.fsrc = as1(m->fsrc);
    q.Kv = as1(m->Kv);
    q.Kv32 = as1(m->Kv32);
    q.mask = as1(m->mask);
    q.dinv = as1(m->dinv);
    q.binv = as1(m->binv);
    q.dinv32 = as1(m->dinv32);
    q.binv32 = as1(m->binv32);
    q.dF = as1(m->dF);
    q.estat = as1(m->estat);
    return q;
}

// The stiffness kernel a model's launch would use: 0..11 the exact kernel
// (shape * 3 + tiling), 100 + shape the Gram-form matrix-core kernel, 200 /
// 201 the HEX20 matrix-core kernel with 27 / 8 IPs.
int ke_key(const amr_model *m)
{
    const int sh = shape_index(m);
    if (m->ke_path == AMR_KE_MFMA && m->all_elastic) return 100 + sh;
    if (m->ke_path == AMR_KE_MFMA && m->etype == AMR_HEX20) return m->nips == 27 ? 200 : 201;
    return sh * 3 + ke_variant(m, sh);
}

int up_key(const amr_model *m)
{
    const int sh = shape_index(m);
    return sh * 3 + up_variant(m, sh);
}

// One upload of the jobs and every launch's start table; returns the device
// addresses (offsets into the staging buffer)
struct BatchUpload {
    size_t off = 0;
    std::vector<std::pair<size_t, std::vector<char>>> parts;
    template <typename T>
    size_t add(const std::vector<T> &v)
    {
        const size_t o = off;
        std::vector<char> b(v.size() * sizeof(T));
        if (!v.empty()) std::memcpy(b.data(), v.data(), b.size());
        parts.emplace_back(o, std::move(b));
        off = (off + v.size() * sizeof(T) + 15) & ~(size_t)15;
        return o;
    }
    int send(AmrStaging &stg, hipStream_t s)
    {
        if (int rc = amr_staging_reserve(stg, s, std::max<size_t>(off, 16))) return rc;
        for (const auto &p : parts)
            if (!p.second.empty()) std::memcpy(stg.h + p.first, p.second.data(), p.second.size());
        AMR_HIP(hipMemcpyAsync(stg.d, stg.h, off, hipMemcpyHostToDevice, s));
        AMR_HIP(hipEventRecord(stg.ev, s));
        stg.pending = true;
        return AMR_OK;
    }
};

// start table of a launch over `idx` (indices into the jobs) with `wgs(j)`
// workgroups for job j (the jobs of the launch are copied in that order):
// the offsets, then the job of every 1,024th workgroup (job_of)
template <typename F>
std::vector<int64_t> starts(const std::vector<int> &idx, F wgs)
{
    const size_t n = idx.size();
    std::vector<int64_t> st(n + 1, 0);
    for (size_t i = 0; i < n; i++) st[i + 1] = st[i] + wgs(idx[i]);
    const int64_t total = st[n], nh = (total >> HINT_SHIFT) + 1;
    size_t j = 0;
    for (int64_t k = 0; k < nh; k++) {
        const int64_t b = k << HINT_SHIFT;
        while (j + 1 < n && st[j + 1] <= b) j++;
        st.push_back((int64_t)j);
    }
    return st;
}

int launch_ke_batch(int key, const ElemJob *J, const int64_t *st, int n, unsigned grid, hipStream_t s)
{
#define KEB(NN, NQ, A) hipLaunchKernelGGL((ke_batch_kernel<NN, NQ, A>), dim3(grid), dim3(64), 0, s, J, st, n)
    switch (key) {
    case 0: KEB(8, 8, 1); break;
    case 1: KEB(8, 8, 2); break;
    case 2: KEB(8, 8, 4); break;
    case 3: KEB(20, 27, 7); break;
    case 4: KEB(20, 27, 10); break;
    case 5: KEB(20, 27, 20); break;
    case 6: KEB(20, 8, 7); break;
    case 7: KEB(20, 8, 10); break;
    case 8: KEB(20, 8, 20); break;
    case 9: KEB(10, 4, 2); break;
    case 10: KEB(10, 4, 5); break;
    case 11: KEB(10, 4, 10); break;
    case 100: hipLaunchKernelGGL((ke_gram_batch_kernel<8, 8>), dim3(grid), dim3(64), 0, s, J, st, n); break;
    case 101: hipLaunchKernelGGL((ke_gram_batch_kernel<20, 27>), dim3(grid), dim3(64), 0, s, J, st, n); break;
    case 102: hipLaunchKernelGGL((ke_gram_batch_kernel<20, 8>), dim3(grid), dim3(64), 0, s, J, st, n); break;
    case 103: hipLaunchKernelGGL((ke_gram_batch_kernel<10, 4>), dim3(grid), dim3(64), 0, s, J, st, n); break;
    case 200: hipLaunchKernelGGL((ke_mfma_hex20_batch_kernel<27, 9>), dim3(grid), dim3(64), 0, s, J, st, n); break;
    case 201: hipLaunchKernelGGL((ke_mfma_hex20_batch_kernel<8, 8>), dim3(grid), dim3(64), 0, s, J, st, n); break;
    default: amr_set_error("batched element stiffness: unknown kernel"); return AMR_ERR_ARG;
    }
#undef KEB
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

int launch_update_batch(int key, const ElemJob *J, const int64_t *st, int n, unsigned grid, hipStream_t s)
{
#define UPB(NN, NQ, E) hipLaunchKernelGGL((update_batch_kernel<NN, NQ, E>), dim3(grid), dim3(256), 0, s, J, st, n)
    switch (key) {
    case 0: UPB(8, 8, 32); break;
    case 1: UPB(8, 8, 16); break;
    case 2: UPB(8, 8, 8); break;
    case 3: UPB(20, 27, 9); break;
    case 4: UPB(20, 27, 4); break;
    case 5: UPB(20, 27, 2); break;
    case 6: UPB(20, 8, 16); break;
    case 7: UPB(20, 8, 8); break;
    case 8: UPB(20, 8, 32); break;
    case 9: UPB(10, 4, 32); break;
    case 10: UPB(10, 4, 16); break;
    case 11: UPB(10, 4, 64); break;
    default: amr_set_error("batched update: unknown kernel"); return AMR_ERR_ARG;
    }
#undef UPB
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

// workgroups of a model's update launch with tiling key k
int64_t update_wgs(int key, int64_t ne)
{
    constexpr int EPW[12] = {32, 16, 8, 9, 4, 2, 16, 8, 32, 32, 16, 64};
    return (ne + EPW[key] - 1) / EPW[key];
}

}  // namespace

int amr_batch_mount_K(amr_model *const *ms, int n, hipStream_t s, AmrStaging &stg)
{
    if (n <= 0) return AMR_OK;
    std::vector<ElemJob> jobs(n);
    for (int j = 0; j < n; j++) {
        if (shape_index(ms[j]) < 0) {
            amr_set_error("unknown element type");
            return AMR_ERR_ARG;
        }
        jobs[j] = elem_job(ms[j]);
    }
    // the stiffness launches: one per kernel the models use
    std::map<int, std::vector<int>> groups;
    for (int j = 0; j < n; j++)
        if (ms[j]->ne > 0) groups[ke_key(ms[j])].push_back(j);
    std::vector<int> all, withk;
    for (int j = 0; j < n; j++) {
        all.push_back(j);
        if (jobs[j].nblk > 0) withk.push_back(j);
    }
    BatchUpload up;
    const size_t o_all = up.add(jobs);
    struct L {
        int key;
        size_t o_jobs, o_st;
        int n;
        int64_t wgs;
    };
    std::vector<L> ke;
    auto add_launch = [&](const std::vector<int> &idx, auto wgs, int key) {
        std::vector<ElemJob> jj;
        for (int j : idx) jj.push_back(jobs[j]);
        const std::vector<int64_t> st = starts(idx, wgs);
        L l{key, up.add(jj), 0, (int)idx.size(), st[idx.size()]};
        l.o_st = up.add(st);
        return l;
    };
    for (auto &g : groups) ke.push_back(add_launch(g.second, [&](int j) { return jobs[j].ne; }, g.first));
    const L as = add_launch(withk, [&](int j) { return (jobs[j].nblk + 255) / 256; }, 0);
    const L jc = add_launch(withk, [&](int j) { return (jobs[j].nnodes * 3 + 255) / 256; }, 0);
    const L bj = add_launch(withk, [&](int j) { return (jobs[j].nnodes + 255) / 256; }, 0);
    if (int rc = up.send(stg, s)) return rc;
    auto dj = [&](size_t o) { return (const ElemJob *)(stg.d + o); };
    auto ds = [&](size_t o) { return (const int64_t *)(stg.d + o); };
    hipLaunchKernelGGL(err_reset_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dj(o_all), n);
    AMR_HIP(hipGetLastError());
    for (const L &l : ke)
        if (l.wgs > 0)
            if (int rc = launch_ke_batch(l.key, dj(l.o_jobs), ds(l.o_st), l.n, (unsigned)l.wgs, s)) return rc;
    if (as.wgs > 0) {
        hipLaunchKernelGGL(assemble_batch_kernel, dim3((unsigned)as.wgs), dim3(256), 0, s, dj(as.o_jobs),
                           ds(as.o_st), as.n);
        hipLaunchKernelGGL(jacobi_batch_kernel, dim3((unsigned)jc.wgs), dim3(256), 0, s, dj(jc.o_jobs), ds(jc.o_st),
                           jc.n);
        hipLaunchKernelGGL(block_jacobi_batch_kernel, dim3((unsigned)bj.wgs), dim3(256), 0, s, dj(bj.o_jobs),
                           ds(bj.o_st), bj.n);
        AMR_HIP(hipGetLastError());
    }
    for (int j = 0; j < n; j++) ms[j]->k_version++;
    return AMR_OK;
}

int amr_batch_update(amr_model *const *ms, const double *const *dU, double *const *dFin, int n, hipStream_t s,
                     AmrStaging &stg)
{
    if (n <= 0) return AMR_OK;
    std::vector<ElemJob> jobs(n);
    std::map<int, std::vector<int>> groups;
    std::vector<int> all;
    for (int j = 0; j < n; j++) {
        if (shape_index(ms[j]) < 0) {
            amr_set_error("unknown element type");
            return AMR_ERR_ARG;
        }
        jobs[j] = elem_job(ms[j]);
        jobs[j].dU = as1(dU[j]);
        jobs[j].dFin = as1(dFin[j]);
        if (ms[j]->ne > 0) groups[up_key(ms[j])].push_back(j);
        all.push_back(j);
    }
    BatchUpload up;
    struct L {
        int key;
        size_t o_jobs, o_st;
        int n;
        int64_t wgs;
    };
    auto add_launch = [&](const std::vector<int> &idx, auto wgs, int key) {
        std::vector<ElemJob> jj;
        for (int j : idx) jj.push_back(jobs[j]);
        const std::vector<int64_t> st = starts(idx, wgs);
        L l{key, up.add(jj), 0, (int)idx.size(), st[idx.size()]};
        l.o_st = up.add(st);
        return l;
    };
    std::vector<L> ul;
    for (auto &g : groups) {
        const int key = g.first;
        ul.push_back(add_launch(g.second, [&](int j) { return update_wgs(key, jobs[j].ne); }, key));
    }
    const L fg = add_launch(all, [&](int j) { return (jobs[j].nnodes * 3 + 255) / 256; }, 0);
    if (int rc = up.send(stg, s)) return rc;
    auto dj = [&](size_t o) { return (const ElemJob *)(stg.d + o); };
    auto ds = [&](size_t o) { return (const int64_t *)(stg.d + o); };
    for (const L &l : ul)
        if (l.wgs > 0)
            if (int rc = launch_update_batch(l.key, dj(l.o_jobs), ds(l.o_st), l.n, (unsigned)l.wgs, s)) return rc;
    if (fg.wgs > 0) {
        hipLaunchKernelGGL(force_gather_batch_kernel, dim3((unsigned)fg.wgs), dim3(256), 0, s, dj(fg.o_jobs),
                           ds(fg.o_st), fg.n);
        AMR_HIP(hipGetLastError());
    }
    return AMR_OK;
}
