/*
 * CPU oracle for the Amaru.jl FE assembly / constitutive path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (the HIP library) never links or calls it.
 *
 * It restates, in plain C, the arithmetic of:
 *   - shape derivatives HEX8/HEX20/TET10 (src/shape/solids3d.jl:124-191,
 *     472-504, 584-691) and the quadrature tables HEX_IP8, HEX_IP27, TET_IP4
 *     (src/shape/quadrature.jl:110-114, 167-206)
 *   - elem_stiffness(::MechSolid) and setB (src/mech/elem/mech-solid.jl:82-166)
 *   - update_elem!(::MechSolid) (src/mech/elem/mech-solid.jl:243-279)
 *   - calcDe / calcD / update_state! for LinearElastic
 *     (src/mech/mat/linear-elastic.jl:98-137), VonMises 3D
 *     (src/mech/mat/von-mises.jl:104-156), DruckerPrager
 *     (src/mech/mat/drucker-prager.jl:76-149); tensor helpers dev/J2
 *     (src/tools/tensors.jl:7-27)
 *   - mount_K's triplet rule: drop |Ke_ij| < eps, then sum duplicates in
 *     element (insertion) order (src/mech/mech-solver.jl:78-110)
 *
 * Floating-point operation order.  Julia evaluates the small products with
 * BLAS, whose summation order is not specified, so no CPU restatement can
 * reproduce its last bits.  This file fixes one explicit order (every dot
 * product is an fma chain in ascending index order starting from 0; the
 * scale-and-accumulate steps K += coef*t and dF += coef*t are one fma each;
 * divisions stay divisions), and the HIP kernels follow the same order, so
 * device results can be compared to these bit for bit.  Build with
 * -ffp-contract=off so the compiler adds no fusions of its own.
 *
 * State layout per integration point (14 doubles): sig[6] eps[6] epa dlam,
 * Mandel notation (shear components carry sqrt(2)).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NSTATE 14
#define ORA_HEX8 1
#define ORA_HEX20 2
#define ORA_TET10 3
#define MAT_LE 0
#define MAT_VM 1
#define MAT_DP 2

static const double SR2 = 1.4142135623730951;  /* sqrt(2.0) */

/* ---------------------------------------------------------------- shapes */

int ora_nnodes(int etype)
{
    switch (etype) {
    case ORA_HEX8: return 8;
    case ORA_HEX20: return 20;
    case ORA_TET10: return 10;
    }
    return -1;
}

/* Natural coordinates of HEX20 nodes (node numbering of the reference's
 * HEX20, src/shape/solids3d.jl:557-580).  The first 8 rows are the HEX8
 * corners in HEX8 order. */
static const double HEX_NODES[20][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
    {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1},
    {0, -1, -1},  {1, 0, -1},  {0, 1, -1}, {-1, 0, -1},
    {0, -1, 1},   {1, 0, 1},   {0, 1, 1},  {-1, 0, 1},
    {-1, -1, 0},  {1, -1, 0},  {1, 1, 0},  {-1, 1, 0}};

/* dN/dR for HEX8: N_a = (1+r ra)(1+s sa)(1+t ta)/8 */
static void deriv_hex8(const double *R, double *D)
{
    double r = R[0], s = R[1], t = R[2];
    for (int a = 0; a < 8; a++) {
        double ra = HEX_NODES[a][0], sa = HEX_NODES[a][1], ta = HEX_NODES[a][2];
        double fr = 1.0 + ra * r, fs = 1.0 + sa * s, ft = 1.0 + ta * t;
        D[a * 3 + 0] = 0.125 * ra * fs * ft;
        D[a * 3 + 1] = 0.125 * sa * fr * ft;
        D[a * 3 + 2] = 0.125 * ta * fr * fs;
    }
}

static void func_hex8(const double *R, double *N)
{
    for (int a = 0; a < 8; a++)
        N[a] = 0.125 * (1.0 + HEX_NODES[a][0] * R[0]) * (1.0 + HEX_NODES[a][1] * R[1]) *
               (1.0 + HEX_NODES[a][2] * R[2]);
}

/* HEX20 serendipity element.  Corners: N = (1+r ra)(1+s sa)(1+t ta)(r ra + s sa + t ta - 2)/8;
 * mid-edge nodes with a zero natural coordinate along one axis:
 * N = (1-x^2)(1+y ya)(1+z za)/4. */
static void deriv_hex20(const double *R, double *D)
{
    double r = R[0], s = R[1], t = R[2];
    for (int a = 0; a < 8; a++) {
        double ra = HEX_NODES[a][0], sa = HEX_NODES[a][1], ta = HEX_NODES[a][2];
        double fr = 1.0 + ra * r, fs = 1.0 + sa * s, ft = 1.0 + ta * t;
        double g = ra * r + sa * s + ta * t - 2.0;
        D[a * 3 + 0] = 0.125 * ra * fs * ft * (g + fr);
        D[a * 3 + 1] = 0.125 * sa * fr * ft * (g + fs);
        D[a * 3 + 2] = 0.125 * ta * fr * fs * (g + ft);
    }
    for (int a = 8; a < 20; a++) {
        double ra = HEX_NODES[a][0], sa = HEX_NODES[a][1], ta = HEX_NODES[a][2];
        if (ra == 0.0) {
            double fs = 1.0 + sa * s, ft = 1.0 + ta * t, q = 1.0 - r * r;
            D[a * 3 + 0] = -0.5 * r * fs * ft;
            D[a * 3 + 1] = 0.25 * q * sa * ft;
            D[a * 3 + 2] = 0.25 * q * fs * ta;
        } else if (sa == 0.0) {
            double fr = 1.0 + ra * r, ft = 1.0 + ta * t, q = 1.0 - s * s;
            D[a * 3 + 0] = 0.25 * q * ra * ft;
            D[a * 3 + 1] = -0.5 * s * fr * ft;
            D[a * 3 + 2] = 0.25 * q * fr * ta;
        } else {
            double fr = 1.0 + ra * r, fs = 1.0 + sa * s, q = 1.0 - t * t;
            D[a * 3 + 0] = 0.25 * q * ra * fs;
            D[a * 3 + 1] = 0.25 * q * fr * sa;
            D[a * 3 + 2] = -0.5 * t * fr * fs;
        }
    }
}

static void func_hex20(const double *R, double *N)
{
    double r = R[0], s = R[1], t = R[2];
    for (int a = 0; a < 8; a++) {
        double ra = HEX_NODES[a][0], sa = HEX_NODES[a][1], ta = HEX_NODES[a][2];
        N[a] = 0.125 * (1.0 + ra * r) * (1.0 + sa * s) * (1.0 + ta * t) *
               (ra * r + sa * s + ta * t - 2.0);
    }
    for (int a = 8; a < 20; a++) {
        double ra = HEX_NODES[a][0], sa = HEX_NODES[a][1], ta = HEX_NODES[a][2];
        if (ra == 0.0)
            N[a] = 0.25 * (1.0 - r * r) * (1.0 + sa * s) * (1.0 + ta * t);
        else if (sa == 0.0)
            N[a] = 0.25 * (1.0 + ra * r) * (1.0 - s * s) * (1.0 + ta * t);
        else
            N[a] = 0.25 * (1.0 + ra * r) * (1.0 + sa * s) * (1.0 - t * t);
    }
}

/* TET10 in volume coordinates L0 = 1-r-s-t, L1 = r, L2 = s, L3 = t.
 * Corners N_i = L_i(2 L_i - 1); edge nodes N = 4 L_i L_j with the edge list
 * (0,1) (1,2) (2,0) (0,3) (1,3) (2,3) (node numbering of the reference's
 * TET10, src/shape/solids3d.jl:108-119). */
static const int TET_EDGES[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

static void tet_L(const double *R, double L[4], double dL[4][3])
{
    L[1] = R[0];
    L[2] = R[1];
    L[3] = R[2];
    L[0] = 1.0 - R[0] - R[1] - R[2];
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 3; k++) dL[i][k] = 0.0;
    dL[0][0] = dL[0][1] = dL[0][2] = -1.0;
    dL[1][0] = dL[2][1] = dL[3][2] = 1.0;
}

static void deriv_tet10(const double *R, double *D)
{
    double L[4], dL[4][3];
    tet_L(R, L, dL);
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 3; k++) D[i * 3 + k] = (4.0 * L[i] - 1.0) * dL[i][k];
    for (int e = 0; e < 6; e++) {
        int i = TET_EDGES[e][0], j = TET_EDGES[e][1];
        for (int k = 0; k < 3; k++) D[(4 + e) * 3 + k] = 4.0 * (dL[i][k] * L[j] + L[i] * dL[j][k]);
    }
}

static void func_tet10(const double *R, double *N)
{
    double L[4], dL[4][3];
    tet_L(R, L, dL);
    for (int i = 0; i < 4; i++) N[i] = L[i] * (2.0 * L[i] - 1.0);
    for (int e = 0; e < 6; e++) N[4 + e] = 4.0 * L[TET_EDGES[e][0]] * L[TET_EDGES[e][1]];
}

void ora_shape_deriv(int etype, const double *R, double *D)
{
    if (etype == ORA_HEX8) deriv_hex8(R, D);
    else if (etype == ORA_HEX20) deriv_hex20(R, D);
    else if (etype == ORA_TET10) deriv_tet10(R, D);
}

void ora_shape_func(int etype, const double *R, double *N)
{
    if (etype == ORA_HEX8) func_hex8(R, N);
    else if (etype == ORA_HEX20) func_hex20(R, N);
    else if (etype == ORA_TET10) func_tet10(R, N);
}

/* Gauss rules, literal values as tabulated by the reference
 * (src/shape/quadrature.jl:110-114, 167-206): 15-significant-digit points. */
static const double G2 = 0.577350269189626;
static const double G3 = 0.774596669241483;
static const double W3[3] = {0.171467764060357, 0.274348422496571, 0.438957475994513};

/* Fills R (n×3) and w (n); returns n, or -1 if the rule is unknown.
 * HEX rules run r fastest, then s, then t. */
int ora_ip_rule(int etype, int nips, double *R, double *w)
{
    if (etype == ORA_HEX8 || etype == ORA_HEX20) {
        if (nips == 0 || nips == 8) {
            int q = 0;
            for (int k = 0; k < 2; k++)
                for (int j = 0; j < 2; j++)
                    for (int i = 0; i < 2; i++, q++) {
                        R[q * 3 + 0] = i ? G2 : -G2;
                        R[q * 3 + 1] = j ? G2 : -G2;
                        R[q * 3 + 2] = k ? G2 : -G2;
                        w[q] = 1.0;
                    }
            return 8;
        }
        if (nips == 27) {
            /* weights w_i w_j w_k as printed: corner 0.171..., edge 0.274...,
             * face 0.438..., centre 0.702331961591221 */
            const double pts[3] = {-G3, 0.0, G3};
            int q = 0;
            for (int k = 0; k < 3; k++)
                for (int j = 0; j < 3; j++)
                    for (int i = 0; i < 3; i++, q++) {
                        R[q * 3 + 0] = pts[i];
                        R[q * 3 + 1] = pts[j];
                        R[q * 3 + 2] = pts[k];
                        int nz = (i == 1) + (j == 1) + (k == 1);
                        w[q] = nz == 3 ? 0.702331961591221 : W3[nz];
                    }
            return 27;
        }
        return -1;
    }
    if (etype == ORA_TET10) {
        if (nips == 0 || nips == 4) {
            const double a = 0.5854101966249685, b = 0.1381966011250105;
            const double P[4][3] = {{a, b, b}, {b, a, b}, {b, b, a}, {b, b, b}};
            for (int q = 0; q < 4; q++) {
                for (int k = 0; k < 3; k++) R[q * 3 + k] = P[q][k];
                w[q] = 0.04166666666666667;
            }
            return 4;
        }
        return -1;
    }
    return -1;
}

/* ------------------------------------------------------------- materials */

/* Elastic Mandel stiffness (src/mech/mat/linear-elastic.jl:110-117). */
void ora_calcDe(double E, double nu, double *D)
{
    double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double a = c * (1.0 - nu), b = c * nu, g = c * (1.0 - 2.0 * nu);
    for (int i = 0; i < 36; i++) D[i] = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) D[i * 6 + j] = (i == j) ? a : b;
    for (int i = 3; i < 6; i++) D[i * 6 + i] = g;
}

/* y = A x for a 6×6 row-major A: fma chain over columns. */
static void mv6(const double *A, const double *x, double *y)
{
    for (int i = 0; i < 6; i++) {
        double acc = 0.0;
        for (int j = 0; j < 6; j++) acc = fma(A[i * 6 + j], x[j], acc);
        y[i] = acc;
    }
}

static double dot6(const double *a, const double *b)
{
    double acc = 0.0;
    for (int i = 0; i < 6; i++) acc = fma(a[i], b[i], acc);
    return acc;
}

/* deviator, as the product with the 6×6 deviatoric projector
 * (src/tools/tensors.jl:10-18) */
static void dev6(const double *x, double *s)
{
    const double d = 2.0 / 3.0, o = -1.0 / 3.0;
    s[0] = fma(o, x[2], fma(o, x[1], d * x[0]));
    s[1] = fma(o, x[2], fma(d, x[1], o * x[0]));
    s[2] = fma(d, x[2], fma(o, x[1], o * x[0]));
    s[3] = x[3];
    s[4] = x[4];
    s[5] = x[5];
}

/* second deviatoric invariant of a Mandel vector (src/tools/tensors.jl:24-27) */
static double J2(const double *x)
{
    double t11 = x[0], t22 = x[1], t33 = x[2];
    double t23 = x[3] / SR2, t13 = x[4] / SR2, t12 = x[5] / SR2;
    double a = t11 - t22, b = t22 - t33, c = t33 - t11;
    return (1.0 / 6.0) * (a * a + b * b + c * c) + t23 * t23 + t13 * t13 + t12 * t12;
}

/* params: LE {E,nu}; VM {E,nu,fy,H}; DP {E,nu,alpha,kappa,H} */
void ora_calcD(int kind, const double *p, const double *st, double *D)
{
    ora_calcDe(p[0], p[1], D);
    if (kind == MAT_VM) {
        /* src/mech/mat/von-mises.jl:112-125 */
        if (st[13] == 0.0) return;
        double s[6], n[6], v[6];
        dev6(st, s);
        double nrm = sqrt(dot6(s, s));
        for (int i = 0; i < 6; i++) n[i] = (sqrt(1.5) * s[i]) / nrm;
        double De[36];
        memcpy(De, D, sizeof De);
        mv6(De, n, v);
        double den = dot6(n, v) - sqrt(1.5) * (-p[3]);
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) D[i * 6 + j] = De[i * 6 + j] - (v[i] * v[j]) / den;
    } else if (kind == MAT_DP) {
        /* src/mech/mat/drucker-prager.jl:86-109 */
        if (st[13] == 0.0) return;
        double alpha = p[2], H = p[4];
        double Nu[6], V[6];
        const double I2[6] = {1, 1, 1, 0, 0, 0};
        if (J2(st) != 0.0) {
            double s[6];
            dev6(st, s);
            double ns = sqrt(dot6(s, s));
            for (int i = 0; i < 6; i++) V[i] = alpha * I2[i] + (s[i] / ns) / SR2;
            double nv = sqrt(dot6(V, V));
            for (int i = 0; i < 6; i++) Nu[i] = V[i] / nv;
        } else {
            for (int i = 0; i < 6; i++) Nu[i] = V[i] = (1.0 / sqrt(3.0)) * I2[i];
        }
        double De[36], a[6], b[6];
        memcpy(De, D, sizeof De);
        mv6(De, Nu, a);
        mv6(De, V, b);
        double den = dot6(b, Nu) + H;
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) D[i * 6 + j] = De[i * 6 + j] - (a[i] * b[j]) / den;
    }
}

/* Returns 0 on success, 1 on a material failure (state untouched except for
 * the fields the reference also writes before failing). */
int ora_update_state(int kind, const double *p, double *st, const double *de, double *dsig)
{
    double De[36], sini[6], str[6], t[6];
    ora_calcDe(p[0], p[1], De);
    for (int i = 0; i < 6; i++) sini[i] = st[i];
    mv6(De, de, t);
    for (int i = 0; i < 6; i++) str[i] = st[i] + t[i];

    if (kind == MAT_LE) {
        /* src/mech/mat/linear-elastic.jl:130-137 */
        for (int i = 0; i < 6; i++) {
            st[6 + i] += de[i];
            st[i] = str[i];
            dsig[i] = t[i];
        }
        return 0;
    }
    if (kind == MAT_VM) {
        /* src/mech/mat/von-mises.jl:104-109, 128-156 */
        double E = p[0], nu = p[1], fy = p[2], H = p[3];
        double j2 = J2(str);
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
            dev6(str, s);
            for (int i = 0; i < 6; i++) s[i] = f * s[i];
            double ns = sqrt(dot6(s, s));
            double c = sqrt(6.0) * G * dl;
            for (int i = 0; i < 6; i++) st[i] = str[i] - (c * s[i]) / ns;
            st[12] += dl;
        }
    } else if (kind == MAT_DP) {
        /* src/mech/mat/drucker-prager.jl:76-83, 112-149 */
        double E = p[0], nu = p[1], alpha = p[2], kappa = p[3], H = p[4];
        double j1 = str[0] + str[1] + str[2];
        double j2 = J2(str);
        double ftr = alpha * j1 + sqrt(j2) - kappa - H * st[12];
        if (ftr < 1e-8) {
            st[13] = 0.0;
            for (int i = 0; i < 6; i++) st[i] = str[i];
        } else {
            double K = E / (3.0 * (1.0 - 2.0 * nu)), G = E / (2.0 * (1.0 + nu));
            double n = 1.0 / sqrt(3.0 * alpha * alpha + 0.5);
            double sj = sqrt(j2);
            /* the branch test reads the previous step's dgamma, as the reference does */
            if (sj - st[13] * n * G > 0.0) {
                double dg = ftr / (9.0 * alpha * alpha * n * K + n * G + H);
                st[13] = dg;
                double j1n = j1 - 9.0 * dg * alpha * n * K;
                double m = 1.0 - dg * n * G / sj;
                double s[6];
                dev6(str, s);
                for (int i = 0; i < 6; i++) st[i] = m * s[i] + (i < 3 ? j1n / 3.0 : 0.0);
            } else {
                double dg = (alpha * j1 - kappa - H * st[12]) / (3.0 * sqrt(3.0) * alpha * K + H);
                st[13] = dg;
                double j1n = j1 - 3.0 * sqrt(3.0) * dg * K;
                for (int i = 0; i < 6; i++) st[i] = (i < 3 ? j1n / 3.0 : 0.0);
            }
            st[12] += st[13];
        }
    } else {
        return 1;
    }
    for (int i = 0; i < 6; i++) {
        st[6 + i] += de[i];
        dsig[i] = st[i] - sini[i];
    }
    return 0;
}

/* ------------------------------------------------------ element geometry */

/* Per-IP geometry: dNdX (nn×3) and coef = detJ*w.  Returns 0, or 1 if the
 * Jacobian determinant is not positive. */
static int ip_geometry(int nn, const double *dNdR, double w, const double *C, double *dNdX, double *coef)
{
    double J[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
            for (int n = 0; n < nn; n++) acc = fma(C[n * 3 + i], dNdR[n * 3 + j], acc);
            J[i][j] = acc;
        }
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    if (!(det > 0.0)) return 1;
    double inv[3][3];
    inv[0][0] = c00 / det;
    inv[1][0] = c01 / det;
    inv[2][0] = c02 / det;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    for (int n = 0; n < nn; n++)
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
            for (int i = 0; i < 3; i++) acc = fma(dNdR[n * 3 + i], inv[i][j], acc);
            dNdX[n * 3 + j] = acc;
        }
    *coef = det * w;
    return 0;
}

/* Dense Mandel B (6 × 3nn, row-major), setB of src/mech/elem/mech-solid.jl:106-118 */
static void build_B(int nn, const double *dNdX, double *B)
{
    int nd = 3 * nn;
    memset(B, 0, sizeof(double) * 6 * nd);
    for (int a = 0; a < nn; a++) {
        double gx = dNdX[a * 3], gy = dNdX[a * 3 + 1], gz = dNdX[a * 3 + 2];
        int c = 3 * a;
        B[0 * nd + c] = gx;
        B[1 * nd + c + 1] = gy;
        B[2 * nd + c + 2] = gz;
        B[3 * nd + c + 1] = gz / SR2;
        B[3 * nd + c + 2] = gy / SR2;
        B[4 * nd + c] = gz / SR2;
        B[4 * nd + c + 2] = gx / SR2;
        B[5 * nd + c] = gy / SR2;
        B[5 * nd + c + 1] = gx / SR2;
    }
}

#define MAXNN 20
#define MAXND 60
#define MAXIP 27

typedef struct {
    int nn, nips;
    double R[MAXIP * 3], w[MAXIP];
    double dNdR[MAXIP][MAXNN * 3];
} rule_t;

static int make_rule(int etype, int nips, rule_t *r)
{
    r->nn = ora_nnodes(etype);
    if (r->nn < 0) return -1;
    r->nips = ora_ip_rule(etype, nips, r->R, r->w);
    if (r->nips < 0) return -1;
    for (int q = 0; q < r->nips; q++) ora_shape_deriv(etype, r->R + 3 * q, r->dNdR[q]);
    return 0;
}

/* dN/dR of every IP of a rule: out is nips × nn × 3 */
int ora_rule_derivs(int etype, int nips, double *out)
{
    rule_t r;
    if (make_rule(etype, nips, &r)) return -1;
    for (int q = 0; q < r.nips; q++) memcpy(out + q * r.nn * 3, r.dNdR[q], sizeof(double) * r.nn * 3);
    return r.nips;
}

static void gather_coords(int nn, const double *X, const int32_t *conn, double *C)
{
    for (int a = 0; a < nn; a++)
        for (int k = 0; k < 3; k++) C[a * 3 + k] = X[(int64_t)conn[a] * 3 + k];
}

/*
 * Element stiffness for ne elements of one type (elem_stiffness,
 * src/mech/elem/mech-solid.jl:124-166).  For each IP:
 *   DB[k][c] = fma-chain_l D[k][l] B[l][c]
 *   t        = fma-chain_k B[k][i] DB[k][j]
 *   K[i][j]  = fma(coef, t, K[i][j])
 * Ke is ne × nd × nd row-major.  states: nips per element × 14.  Returns 0,
 * or 1 + (index of the first element with a non-positive Jacobian).
 */
int64_t ora_elem_stiffness(int etype, int nips, int64_t ne, const double *X, const int32_t *conn,
                           const int32_t *emat, const int32_t *kinds, const double *prms, int prm_stride,
                           const double *states, double *Ke, int nthreads)
{
    rule_t r;
    if (make_rule(etype, nips, &r)) return -1;
    int nn = r.nn, nd = 3 * nn, nq = r.nips;
    int64_t bad = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int64_t e = 0; e < ne; e++) {
        double C[MAXNN * 3], dNdX[MAXNN * 3], B[6 * MAXND], DB[6 * MAXND], D[36], coef;
        double *K = Ke + e * nd * nd;
        memset(K, 0, sizeof(double) * nd * nd);
        gather_coords(nn, X, conn + e * nn, C);
        int m = emat[e];
        for (int q = 0; q < nq; q++) {
            if (ip_geometry(nn, r.dNdR[q], r.w[q], C, dNdX, &coef)) {
#ifdef _OPENMP
#pragma omp critical
#endif
                if (bad == 0 || e + 1 < bad) bad = e + 1;
                break;
            }
            build_B(nn, dNdX, B);
            ora_calcD(kinds[m], prms + (int64_t)m * prm_stride, states + (e * nq + q) * NSTATE, D);
            for (int k = 0; k < 6; k++)
                for (int c = 0; c < nd; c++) {
                    double acc = 0.0;
                    for (int l = 0; l < 6; l++) acc = fma(D[k * 6 + l], B[l * nd + c], acc);
                    DB[k * nd + c] = acc;
                }
            for (int i = 0; i < nd; i++)
                for (int j = 0; j < nd; j++) {
                    double t = 0.0;
                    for (int k = 0; k < 6; k++) t = fma(B[k * nd + i], DB[k * nd + j], t);
                    K[i * nd + j] = fma(coef, t, K[i * nd + j]);
                }
        }
    }
    return bad;
}

/*
 * update_elem! for ne elements (src/mech/elem/mech-solid.jl:243-279):
 *   dU = U[map];  per IP: deps = B dU (fma chain over columns);
 *   (dsig, status) = update_state!(mat, state, deps);  dF += coef*B' dsig.
 * dF is ne × nd.  status[e] = 0 ok, 1 material failure, 2 bad Jacobian.
 * On a failure the element stops at that IP, like the reference.
 */
int ora_update_elems(int etype, int nips, int64_t ne, const double *X, const int32_t *conn,
                     const int32_t *eqmap, const int32_t *emat, const int32_t *kinds, const double *prms,
                     int prm_stride, const double *U, double *states, double *dF, int32_t *status,
                     int nthreads)
{
    rule_t r;
    if (make_rule(etype, nips, &r)) return -1;
    int nn = r.nn, nd = 3 * nn, nq = r.nips;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int64_t e = 0; e < ne; e++) {
        double C[MAXNN * 3], dNdX[MAXNN * 3], B[6 * MAXND], dU[MAXND], coef, de[6], ds[6];
        double *F = dF + e * nd;
        memset(F, 0, sizeof(double) * nd);
        gather_coords(nn, X, conn + e * nn, C);
        for (int c = 0; c < nd; c++) dU[c] = U[eqmap[e * nd + c]];
        int m = emat[e];
        status[e] = 0;
        for (int q = 0; q < nq; q++) {
            if (ip_geometry(nn, r.dNdR[q], r.w[q], C, dNdX, &coef)) {
                status[e] = 2;
                break;
            }
            build_B(nn, dNdX, B);
            for (int k = 0; k < 6; k++) {
                double acc = 0.0;
                for (int c = 0; c < nd; c++) acc = fma(B[k * nd + c], dU[c], acc);
                de[k] = acc;
            }
            if (ora_update_state(kinds[m], prms + (int64_t)m * prm_stride, states + (e * nq + q) * NSTATE,
                                 de, ds)) {
                status[e] = 1;
                break;
            }
            for (int c = 0; c < nd; c++) {
                double t = 0.0;
                for (int k = 0; k < 6; k++) t = fma(B[k * nd + c], ds[k], t);
                F[c] = fma(coef, t, F[c]);
            }
        }
    }
    return 0;
}

/*
 * ΔFin gather: ΔFin[map] .+= ΔF element by element in element order
 * (src/mech/mech-solver.jl:129-136, single-thread order).
 */
void ora_gather_forces(int64_t ne, int nd, const double *dF, const int32_t *eqmap, int64_t ndofs, double *out)
{
    memset(out, 0, sizeof(double) * ndofs);
    for (int64_t e = 0; e < ne; e++)
        for (int c = 0; c < nd; c++) out[eqmap[e * nd + c]] += dF[e * nd + c];
}

/*
 * Element mass matrices (elem_mass, src/mech/elem/mech-solid.jl:169-205) for
 * ne elements of one type, rho per element, thickness 1.  For each IP:
 *   coef = (rho * detJ) * w            (detJ as in ip_geometry)
 *   M[(a,i)][(b,i)] = M + coef * (N_a * N_b)   (N'N has one nonzero term)
 * entries with i != j stay 0.  Me is ne × nd × nd row-major (dof 3a+i).
 * Returns 0, or 1 + (index of the first element with a non-positive
 * Jacobian).
 */
int64_t ora_elem_mass(int etype, int nips, int64_t ne, const double *X, const int32_t *conn, const double *rho,
                      double *Me)
{
    rule_t r;
    if (make_rule(etype, nips, &r)) return -1;
    const int nn = r.nn, nd = 3 * nn, nq = r.nips;
    double N[MAXIP][MAXNN];
    for (int q = 0; q < nq; q++) ora_shape_func(etype, r.R + 3 * q, N[q]);
    for (int64_t e = 0; e < ne; e++) {
        double C[MAXNN * 3], dNdX[MAXNN * 3], dcoef;
        gather_coords(nn, X, conn + e * nn, C);
        double *M = Me + e * nd * nd;
        memset(M, 0, sizeof(double) * nd * nd);
        for (int q = 0; q < nq; q++) {
            if (ip_geometry(nn, r.dNdR[q], 1.0, C, dNdX, &dcoef)) return e + 1;
            const double coef = rho[e] * dcoef * r.w[q];
            for (int a = 0; a < nn; a++)
                for (int b = 0; b < nn; b++) {
                    const double t = N[q][a] * N[q][b];
                    for (int i = 0; i < 3; i++) M[(a * 3 + i) * nd + b * 3 + i] += coef * t;
                }
        }
    }
    return 0;
}

/* ------------------------------------------------------------- mount_K */

/* Number of triplets that survive the |v| < eps filter. */
int64_t ora_count_triplets(int64_t ne, int nd, const double *Ke)
{
    const double eps = 2.220446049250313e-16;
    int64_t n = 0;
    for (int64_t i = 0; i < ne * nd * nd; i++)
        if (!(fabs(Ke[i]) < eps)) n++;
    return n;
}

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/*
 * mount_K (src/mech/mech-solver.jl:78-110): triplets in element order with
 * |Ke_ij| < eps dropped, then duplicates summed in insertion order
 * (first value, then += each later one).  Output CSR with sorted columns:
 * rowptr[ndofs+1], col[nnz], val[nnz]; col/val must hold ntrip entries
 * (ora_count_triplets).  Returns nnz, or -1 on allocation failure.
 */
int64_t ora_mount_K(int64_t ne, int nd, const double *Ke, const int32_t *eqmap, int64_t ndofs,
                    int64_t *rowptr, int32_t *col, double *val)
{
    const double eps = 2.220446049250313e-16;
    /* counting sort of triplets by row, stable in element order */
    int64_t *cnt = calloc(ndofs + 1, sizeof(int64_t));
    if (!cnt) return -1;
    for (int64_t e = 0; e < ne; e++)
        for (int i = 0; i < nd; i++)
            for (int j = 0; j < nd; j++)
                if (!(fabs(Ke[(e * nd + i) * nd + j]) < eps)) cnt[eqmap[e * nd + i] + 1]++;
    for (int64_t r = 0; r < ndofs; r++) cnt[r + 1] += cnt[r];
    int64_t ntrip = cnt[ndofs];
    int32_t *tc = malloc(sizeof(int32_t) * (ntrip ? ntrip : 1));
    double *tv = malloc(sizeof(double) * (ntrip ? ntrip : 1));
    int64_t *pos = malloc(sizeof(int64_t) * (ndofs + 1));
    int64_t *slot = malloc(sizeof(int64_t) * (ndofs ? ndofs : 1));
    if (!tc || !tv || !pos || !slot) {
        free(cnt); free(tc); free(tv); free(pos); free(slot);
        return -1;
    }
    memcpy(pos, cnt, sizeof(int64_t) * (ndofs + 1));
    for (int64_t e = 0; e < ne; e++)
        for (int i = 0; i < nd; i++)
            for (int j = 0; j < nd; j++) {
                double v = Ke[(e * nd + i) * nd + j];
                if (fabs(v) < eps) continue;
                int64_t p = pos[eqmap[e * nd + i]]++;
                tc[p] = eqmap[e * nd + j];
                tv[p] = v;
            }
    for (int64_t c = 0; c < ndofs; c++) slot[c] = -1;
    int64_t nnz = 0;
    rowptr[0] = 0;
    for (int64_t r = 0; r < ndofs; r++) {
        int64_t start = nnz;
        for (int64_t p = cnt[r]; p < cnt[r + 1]; p++) {
            int32_t c = tc[p];
            if (slot[c] < 0) {
                slot[c] = nnz;
                col[nnz] = c;
                val[nnz] = tv[p];
                nnz++;
            } else {
                val[slot[c]] += tv[p];
            }
        }
        /* sort the row's unique columns; values follow their column */
        int64_t len = nnz - start;
        qsort(col + start, len, sizeof(int32_t), cmp_i32);
        /* slot[] still holds each column's unsorted position */
        {
            double *tmp = malloc(sizeof(double) * (len ? len : 1));
            for (int64_t k = 0; k < len; k++) tmp[k] = val[slot[col[start + k]]];
            for (int64_t k = 0; k < len; k++) {
                val[start + k] = tmp[k];
                slot[col[start + k]] = -1;
            }
            free(tmp);
        }
        rowptr[r + 1] = nnz;
    }
    free(cnt); free(tc); free(tv); free(pos); free(slot);
    return nnz;
}
