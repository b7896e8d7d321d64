// Distributed loads on the device: the per-cell force vectors of
// mech_boundary_forces (tractions tx/ty/tz/tn on facets, line loads on
// edges; src/mech/elem/distributed.jl:75-150) and of the solid body forces
// (wx/wy/wz; :155-210), for many cells at once.
//
// A cell is a facet (ndim 2), an edge (ndim 1) or an element (ndim 3) with
// npts nodes; the caller passes the cell shape's functions N, derivatives D
// and weights w at its integration points.  Per cell and point:
//   J = C' D (3 x ndim), X = C' N, v = f(X, t)
//   facets/edges: measure |J1 x J2| or |J1|; direction e_d, or the unit
//                 normal J1 x J2 / |J1 x J2| (tn);  F_a += ((m w) N_a) Q
//   elements:     measure det J;              F_a,d += ((m w) v) N_a
// The load value f is a postfix program over x, y, z, t (the reference's
// expression values), evaluated on a stack of at most AMR_EXPR_STACK values;
// programs are checked on the host before any launch.
#include <cmath>
#include <vector>

#include "amr_internal.h"

namespace {

constexpr int MAXP = 20;   // nodes of a cell
constexpr int MAXQ = 27;   // integration points

__device__ double eval_expr(const int32_t *ops, const double *consts, int nops, double x, double y, double z,
                            double t)
{
    double st[AMR_EXPR_STACK];
    int sp = 0;
    for (int i = 0; i < nops; i++) {
        const int op = ops[2 * i], arg = ops[2 * i + 1];
        switch (op) {
        case AMR_OP_CONST: st[sp++] = consts[arg]; break;
        case AMR_OP_X: st[sp++] = x; break;
        case AMR_OP_Y: st[sp++] = y; break;
        case AMR_OP_Z: st[sp++] = z; break;
        case AMR_OP_T: st[sp++] = t; break;
        case AMR_OP_NEG: st[sp - 1] = -st[sp - 1]; break;
        case AMR_OP_SQRT: st[sp - 1] = sqrt(st[sp - 1]); break;
        case AMR_OP_ABS: st[sp - 1] = fabs(st[sp - 1]); break;
        case AMR_OP_SIN: st[sp - 1] = sin(st[sp - 1]); break;
        case AMR_OP_COS: st[sp - 1] = cos(st[sp - 1]); break;
        case AMR_OP_EXP: st[sp - 1] = exp(st[sp - 1]); break;
        case AMR_OP_LOG: st[sp - 1] = log(st[sp - 1]); break;
        case AMR_OP_TAN: st[sp - 1] = tan(st[sp - 1]); break;
        case AMR_OP_LOG10: st[sp - 1] = log10(st[sp - 1]); break;
        default: {
            const double b = st[--sp], a = st[sp - 1];
            double r;
            switch (op) {
            case AMR_OP_ADD: r = a + b; break;
            case AMR_OP_SUB: r = a - b; break;
            case AMR_OP_MUL: r = a * b; break;
            case AMR_OP_DIV: r = a / b; break;
            case AMR_OP_POW: r = pow(a, b); break;
            case AMR_OP_MAX: r = fmax(a, b); break;
            case AMR_OP_MIN: r = fmin(a, b); break;
            default: r = trunc((a - fmod(a, b)) / b);  // div: exact remainder, truncated quotient
            }
            st[sp - 1] = r;
        }
        }
    }
    return st[0];
}

__global__ void __launch_bounds__(256) cell_loads_kernel(int64_t ncells, int npts, int ndim, int nq,
                                                         const double *__restrict__ C, const double *__restrict__ N,
                                                         const double *__restrict__ D, const double *__restrict__ w,
                                                         int dir, const int32_t *__restrict__ ops,
                                                         const double *__restrict__ consts, int nops, double t,
                                                         double *__restrict__ F)
{
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= ncells) return;
    const double *Cc = C + c * npts * 3;
    double *Fc = F + c * npts * 3;
    for (int i = 0; i < npts * 3; i++) Fc[i] = 0.0;
    for (int q = 0; q < nq; q++) {
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, X[3] = {0, 0, 0};
        for (int k = 0; k < 3; k++) {
            for (int i = 0; i < ndim; i++) {
                double a = 0.0;
                for (int p = 0; p < npts; p++) a += Cc[p * 3 + k] * D[(q * npts + p) * ndim + i];
                J[k][i] = a;
            }
            double a = 0.0;
            for (int p = 0; p < npts; p++) a += Cc[p * 3 + k] * N[q * npts + p];
            X[k] = a;
        }
        const double v = eval_expr(ops, consts, nops, X[0], X[1], X[2], t);
        if (ndim == 3) {
            const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            const double cv = det * w[q] * v;
            for (int p = 0; p < npts; p++) Fc[p * 3 + dir] += cv * N[q * npts + p];
            continue;
        }
        double Q[3] = {0.0, 0.0, 0.0}, m;
        if (ndim == 2) {
            const double n0 = J[1][0] * J[2][1] - J[2][0] * J[1][1];
            const double n1 = J[2][0] * J[0][1] - J[0][0] * J[2][1];
            const double n2 = J[0][0] * J[1][1] - J[1][0] * J[0][1];
            m = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
            if (dir < 0) {
                Q[0] = v * n0 / m;
                Q[1] = v * n1 / m;
                Q[2] = v * n2 / m;
            }
        } else {
            m = sqrt(J[0][0] * J[0][0] + J[1][0] * J[1][0] + J[2][0] * J[2][0]);
        }
        if (dir >= 0) Q[dir] = v;
        const double mw = m * w[q];
        for (int p = 0; p < npts; p++) {
            const double s = mw * N[q * npts + p];
            for (int k = 0; k < 3; k++) Fc[p * 3 + k] += s * Q[k];
        }
    }
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

extern "C" int amr_check_expr(const int32_t *ops, int nops, int nconsts)
{
    if (!ops || nops <= 0) {
        amr_set_error("load expression: empty program");
        return AMR_ERR_ARG;
    }
    int sp = 0;
    for (int i = 0; i < nops; i++) {
        const int op = ops[2 * i], arg = ops[2 * i + 1];
        if (op == AMR_OP_CONST) {
            if (arg < 0 || arg >= nconsts) {
                amr_set_error("load expression: constant index out of range");
                return AMR_ERR_ARG;
            }
            sp++;
        } else if (op >= AMR_OP_X && op <= AMR_OP_T) {
            sp++;
        } else if (op >= AMR_OP_NEG && op <= AMR_OP_LOG10) {
            if (sp < 1) {
                amr_set_error("load expression: stack underflow");
                return AMR_ERR_ARG;
            }
        } else if (op >= AMR_OP_ADD && op <= AMR_OP_IDIV) {
            if (sp < 2) {
                amr_set_error("load expression: stack underflow");
                return AMR_ERR_ARG;
            }
            sp--;
        } else {
            amr_set_error("load expression: unknown operation " + std::to_string(op));
            return AMR_ERR_ARG;
        }
        if (sp > AMR_EXPR_STACK) {
            amr_set_error("load expression: deeper than " + std::to_string(AMR_EXPR_STACK) + " values");
            return AMR_ERR_ARG;
        }
    }
    if (sp != 1) {
        amr_set_error("load expression: the program leaves " + std::to_string(sp) + " values");
        return AMR_ERR_ARG;
    }
    return AMR_OK;
}

extern "C" int amr_cell_loads(int device, int64_t ncells, int npts, int ndim, int nq, const double *C,
                              const double *N, const double *D, const double *w, int dir, const int32_t *ops,
                              const double *consts, int nops, int nconsts, double t, double *F)
{
    if (ncells < 0 || npts < 1 || npts > MAXP || ndim < 1 || ndim > 3 || nq < 1 || nq > MAXQ || !N || !D || !w ||
        (ncells > 0 && (!C || !F)) || dir < -1 || dir > 2 || (dir == -1 && ndim != 2) || nconsts < 0 ||
        (nconsts > 0 && !consts)) {
        amr_set_error("amr_cell_loads: bad sizes, direction or null pointer");
        return AMR_ERR_ARG;
    }
    const int rc0 = amr_check_expr(ops, nops, nconsts);
    if (rc0) return rc0;
    if (ncells == 0) return AMR_OK;
    AMR_HIP(hipSetDevice(device));
    const size_t nc = (size_t)ncells * npts * 3;
    DevBufs bufs;
    double *dC = nullptr, *dN = nullptr, *dD = nullptr, *dw = nullptr, *dK = nullptr, *dF = nullptr;
    int32_t *dops = nullptr;
    int rc = dmalloc(&dC, nc);
    bufs.p.push_back(dC);
    if (!rc) rc = dmalloc(&dN, (size_t)nq * npts);
    bufs.p.push_back(dN);
    if (!rc) rc = dmalloc(&dD, (size_t)nq * npts * ndim);
    bufs.p.push_back(dD);
    if (!rc) rc = dmalloc(&dw, (size_t)nq);
    bufs.p.push_back(dw);
    if (!rc) rc = dmalloc(&dK, (size_t)nconsts);
    bufs.p.push_back(dK);
    if (!rc) rc = dmalloc(&dops, (size_t)nops * 2);
    bufs.p.push_back(dops);
    if (!rc) rc = dmalloc(&dF, nc);
    bufs.p.push_back(dF);
    if (rc) return rc;
    hipStream_t s = nullptr;
    AMR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    AMR_HIP(hipMemcpyAsync(dC, C, sizeof(double) * nc, hipMemcpyHostToDevice, s));
    AMR_HIP(hipMemcpyAsync(dN, N, sizeof(double) * nq * npts, hipMemcpyHostToDevice, s));
    AMR_HIP(hipMemcpyAsync(dD, D, sizeof(double) * nq * npts * ndim, hipMemcpyHostToDevice, s));
    AMR_HIP(hipMemcpyAsync(dw, w, sizeof(double) * nq, hipMemcpyHostToDevice, s));
    if (nconsts) AMR_HIP(hipMemcpyAsync(dK, consts, sizeof(double) * nconsts, hipMemcpyHostToDevice, s));
    AMR_HIP(hipMemcpyAsync(dops, ops, sizeof(int32_t) * nops * 2, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(cell_loads_kernel, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, s, ncells, npts, ndim,
                       nq, dC, dN, dD, dw, dir, dops, dK, nops, t, dF);
    AMR_HIP(hipGetLastError());
    AMR_HIP(hipMemcpyAsync(F, dF, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
    AMR_HIP(hipStreamSynchronize(s));
    return AMR_OK;
}
