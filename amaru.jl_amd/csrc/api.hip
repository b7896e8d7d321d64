// C ABI of libamaru_hip.so (declared in include/amaru_hip.h).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <map>
#include <unordered_map>

#include "amr_internal.h"

#include <rocprofiler-sdk-roctx/roctx.h>

static thread_local std::string g_err;

void amr_set_error(const std::string &msg) { g_err = msg; }

extern "C" const char *amr_last_error(void) { return g_err.c_str(); }

extern "C" int amr_device_count(int *n)
{
    if (!n) return AMR_ERR_ARG;
    AMR_HIP(hipGetDeviceCount(n));
    return AMR_OK;
}

// ------------------------------------------------ shapes and quadrature
// Natural coordinates of the HEX20 nodes (first 8: HEX8 corners), in the
// reference's node numbering (src/shape/solids3d.jl:459-468, 557-580).
static const double kHexNodes[20][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1},
    {0, -1, -1},  {1, 0, -1},  {0, 1, -1}, {-1, 0, -1}, {0, -1, 1},  {1, 0, 1},  {0, 1, 1}, {-1, 0, 1},
    {-1, -1, 0},  {1, -1, 0},  {1, 1, 0},  {-1, 1, 0}};
static const int kTetEdges[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

// dN/dR at R (nn x 3).  HEX8 N_a = (1+r ra)(1+s sa)(1+t ta)/8
// (src/shape/solids3d.jl:472-504); HEX20 serendipity (:584-691); TET10 in
// volume coordinates (:124-191).
static void shape_deriv(int etype, const double *R, double *D)
{
    const double r = R[0], s = R[1], t = R[2];
    if (etype == AMR_HEX8) {
        for (int a = 0; a < 8; a++) {
            const double ra = kHexNodes[a][0], sa = kHexNodes[a][1], ta = kHexNodes[a][2];
            const double fr = 1.0 + ra * r, fs = 1.0 + sa * s, ft = 1.0 + ta * t;
            D[a * 3 + 0] = 0.125 * ra * fs * ft;
            D[a * 3 + 1] = 0.125 * sa * fr * ft;
            D[a * 3 + 2] = 0.125 * ta * fr * fs;
        }
    } else if (etype == AMR_HEX20) {
        for (int a = 0; a < 8; a++) {
            const double ra = kHexNodes[a][0], sa = kHexNodes[a][1], ta = kHexNodes[a][2];
            const double fr = 1.0 + ra * r, fs = 1.0 + sa * s, ft = 1.0 + ta * t;
            const double g = ra * r + sa * s + ta * t - 2.0;
            D[a * 3 + 0] = 0.125 * ra * fs * ft * (g + fr);
            D[a * 3 + 1] = 0.125 * sa * fr * ft * (g + fs);
            D[a * 3 + 2] = 0.125 * ta * fr * fs * (g + ft);
        }
        for (int a = 8; a < 20; a++) {
            const double ra = kHexNodes[a][0], sa = kHexNodes[a][1], ta = kHexNodes[a][2];
            if (ra == 0.0) {
                const double fs = 1.0 + sa * s, ft = 1.0 + ta * t, q = 1.0 - r * r;
                D[a * 3 + 0] = -0.5 * r * fs * ft;
                D[a * 3 + 1] = 0.25 * q * sa * ft;
                D[a * 3 + 2] = 0.25 * q * fs * ta;
            } else if (sa == 0.0) {
                const double fr = 1.0 + ra * r, ft = 1.0 + ta * t, q = 1.0 - s * s;
                D[a * 3 + 0] = 0.25 * q * ra * ft;
                D[a * 3 + 1] = -0.5 * s * fr * ft;
                D[a * 3 + 2] = 0.25 * q * fr * ta;
            } else {
                const double fr = 1.0 + ra * r, fs = 1.0 + sa * s, q = 1.0 - t * t;
                D[a * 3 + 0] = 0.25 * q * ra * fs;
                D[a * 3 + 1] = 0.25 * q * fr * sa;
                D[a * 3 + 2] = -0.5 * t * fr * fs;
            }
        }
    } else {
        double L[4] = {1.0 - r - s - t, r, s, t};
        double dL[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 3; k++) D[i * 3 + k] = (4.0 * L[i] - 1.0) * dL[i][k];
        for (int e = 0; e < 6; e++) {
            const int i = kTetEdges[e][0], j = kTetEdges[e][1];
            for (int k = 0; k < 3; k++) D[(4 + e) * 3 + k] = 4.0 * (dL[i][k] * L[j] + L[i] * dL[j][k]);
        }
    }
}

// N at R (nn): the functions whose derivatives shape_deriv returns
// (src/shape/solids3d.jl:124-146, 472-484, 584-613).
static void shape_func(int etype, const double *R, double *N)
{
    const double r = R[0], s = R[1], t = R[2];
    if (etype == AMR_HEX8 || etype == AMR_HEX20) {
        for (int a = 0; a < 8; a++) {
            const double ra = kHexNodes[a][0], sa = kHexNodes[a][1], ta = kHexNodes[a][2];
            const double f = 0.125 * (1.0 + ra * r) * (1.0 + sa * s) * (1.0 + ta * t);
            N[a] = etype == AMR_HEX8 ? f : f * (ra * r + sa * s + ta * t - 2.0);
        }
        if (etype == AMR_HEX8) return;
        for (int a = 8; a < 20; a++) {
            const double ra = kHexNodes[a][0], sa = kHexNodes[a][1], ta = kHexNodes[a][2];
            if (ra == 0.0)
                N[a] = 0.25 * (1.0 - r * r) * (1.0 + sa * s) * (1.0 + ta * t);
            else if (sa == 0.0)
                N[a] = 0.25 * (1.0 + ra * r) * (1.0 - s * s) * (1.0 + ta * t);
            else
                N[a] = 0.25 * (1.0 + ra * r) * (1.0 + sa * s) * (1.0 - t * t);
        }
    } else {
        const double L[4] = {1.0 - r - s - t, r, s, t};
        for (int i = 0; i < 4; i++) N[i] = L[i] * (2.0 * L[i] - 1.0);
        for (int e = 0; e < 6; e++) N[4 + e] = 4.0 * L[kTetEdges[e][0]] * L[kTetEdges[e][1]];
    }
}

// IP rules, as tabulated by the reference (src/shape/quadrature.jl:110-114,
// 167-206): HEX 2x2x2 and 3x3x3 (r fastest), TET 4-point.
static int ip_rule(int etype, int nips, double *R, double *w)
{
    const double g2 = 0.577350269189626, g3 = 0.774596669241483;
    const double w3[4] = {0.171467764060357, 0.274348422496571, 0.438957475994513, 0.702331961591221};
    if (etype == AMR_HEX8 || etype == AMR_HEX20) {
        if (nips == 0 || nips == 8) {
            int q = 0;
            for (int k = 0; k < 2; k++)
                for (int j = 0; j < 2; j++)
                    for (int i = 0; i < 2; i++, q++) {
                        R[q * 3] = i ? g2 : -g2;
                        R[q * 3 + 1] = j ? g2 : -g2;
                        R[q * 3 + 2] = k ? g2 : -g2;
                        w[q] = 1.0;
                    }
            return 8;
        }
        if (nips == 27) {
            const double pts[3] = {-g3, 0.0, g3};
            int q = 0;
            for (int k = 0; k < 3; k++)
                for (int j = 0; j < 3; j++)
                    for (int i = 0; i < 3; i++, q++) {
                        R[q * 3] = pts[i];
                        R[q * 3 + 1] = pts[j];
                        R[q * 3 + 2] = pts[k];
                        w[q] = w3[(i == 1) + (j == 1) + (k == 1)];
                    }
            return 27;
        }
        return -1;
    }
    if (etype == AMR_TET10 && (nips == 0 || nips == 4)) {
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

// A roctx range over an entry point, named after the reference call it
// implements, so that a profile splits a step by phase (no cost without a
// profiler attached).
struct PhaseRange {
    explicit PhaseRange(const char *name) { roctxRangePushA(name); }
    ~PhaseRange() { roctxRangePop(); }
    PhaseRange(const PhaseRange &) = delete;
    PhaseRange &operator=(const PhaseRange &) = delete;
};

static int nn_of(int etype)
{
    return etype == AMR_HEX8 ? 8 : etype == AMR_HEX20 ? 20 : etype == AMR_TET10 ? 10 : -1;
}

// ------------------------------------------------------------- helpers

// Device memory of models.  A released block is kept in a per-device cache
// and handed to the next request of about its size (best fit, at most 25%
// larger, zeroed on the model's stream), so that creating and destroying
// models in a serving loop does not call hipMalloc or hipFree (hipFree waits
// for the whole device).  Blocks
// are released only after the streams that used them are synchronised
// (free_model, amr_pool_detach_model).  The cache holds at most
// AMR_ARENA_BYTES (environment, default 32 GiB) per device; beyond that, and
// when hipMalloc runs out of memory, cached blocks are freed.
namespace {
struct Arena {
    std::multimap<size_t, void *> blocks;  // size -> cached block
    size_t bytes = 0;
};
std::mutex g_dmem_mu;
struct LiveBlock {
    size_t size;     // bytes of the block
    size_t charged;  // bytes the model asked for (its dev_bytes)
    int device;
};
std::unordered_map<void *, LiveBlock> g_dmem_live;
std::unordered_map<int, Arena> g_arena;

size_t arena_cap()
{
    static const size_t cap = [] {
        const char *e = getenv("AMR_ARENA_BYTES");
        return e ? (size_t)strtoull(e, nullptr, 10) : (size_t)32 << 30;
    }();
    return cap;
}

size_t arena_round(size_t b)
{
    if (b <= 65536) {
        size_t r = 256;
        while (r < b) r <<= 1;
        return r;
    }
    return (b + 65535) & ~(size_t)65535;
}

// frees the cached blocks of a device (lock held, device current)
void arena_trim_locked(int device)
{
    Arena &A = g_arena[device];
    for (auto &kv : A.blocks) (void)hipFree(kv.second);
    A.blocks.clear();
    A.bytes = 0;
}
}  // namespace

int amr_dalloc(amr_model *m, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 1;
    const size_t need = arena_round(bytes);
    const int dev = m->device;
    {
        std::lock_guard<std::mutex> lk(g_dmem_mu);
        Arena &A = g_arena[dev];
        auto it = A.blocks.lower_bound(need);
        if (it != A.blocks.end() && it->first <= need + need / 4) {
            const size_t sz = it->first;
            *p = it->second;
            A.blocks.erase(it);
            A.bytes -= sz;
            g_dmem_live[*p] = {sz, bytes, dev};
            m->dev_bytes += (int64_t)bytes;
        } else {
            *p = nullptr;
        }
    }
    if (*p) {
        // a reused block starts zeroed, as fresh device memory does
        AMR_HIP(hipMemsetAsync(*p, 0, need, m->stream));
        return AMR_OK;
    }
    hipError_t e = hipMalloc(p, need);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(g_dmem_mu);
            arena_trim_locked(dev);
        }
        e = hipMalloc(p, need);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        amr_set_error(std::string("hipMalloc of ") + std::to_string(need) + " bytes failed: " + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? AMR_ERR_NOMEM : AMR_ERR_HIP;
    }
    {
        std::lock_guard<std::mutex> lk(g_dmem_mu);
        g_dmem_live[*p] = {need, bytes, dev};
    }
    m->dev_bytes += (int64_t)bytes;
    return AMR_OK;
}

void amr_release(amr_model *m, void *p)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_dmem_mu);
    auto it = g_dmem_live.find(p);
    if (it == g_dmem_live.end()) {
        (void)hipFree(p);
        return;
    }
    const size_t sz = it->second.size;
    const int dev = it->second.device;
    if (m) m->dev_bytes -= (int64_t)it->second.charged;
    g_dmem_live.erase(it);
    Arena &A = g_arena[dev];
    if (A.bytes + sz <= arena_cap()) {
        A.blocks.insert({sz, p});
        A.bytes += sz;
    } else {
        (void)hipFree(p);
    }
}

extern "C" int amr_mem_trim(int device)
{
    AMR_HIP(hipSetDevice(device));
    std::lock_guard<std::mutex> lk(g_dmem_mu);
    arena_trim_locked(device);
    return AMR_OK;
}

// Device allocation charged to the model's byte count.
template <typename T>
static int dalloc(amr_model *m, T **p, size_t n)
{
    return amr_dalloc(m, (void **)p, (n ? n : 1) * sizeof(T));
}

template <typename T>
static int h2d(T *d, const T *h, size_t n, hipStream_t s)
{
    if (n == 0) return AMR_OK;
    AMR_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    return AMR_OK;
}

template <typename T>
static int d2h(T *h, const T *d, size_t n, hipStream_t s)
{
    if (n == 0) return AMR_OK;
    AMR_HIP(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, s));
    return AMR_OK;
}

#define TRY(x)                 \
    do {                       \
        int _rc = (x);         \
        if (_rc) return _rc;   \
    } while (0)

// Pinned host blocks kept for reuse by later models (pinning memory costs
// about as much as the small copies it serves): best fit within twice the
// request, at most 256 MiB kept.
static std::mutex g_hpin_mu;
static std::multimap<size_t, void *> g_hpin_free;
static size_t g_hpin_kept = 0;

static int pinned_get(size_t bytes, void **out, size_t *got)
{
    {
        std::lock_guard<std::mutex> lk(g_hpin_mu);
        auto it = g_hpin_free.lower_bound(bytes);
        if (it != g_hpin_free.end() && it->first <= 2 * bytes) {
            *out = it->second;
            *got = it->first;
            g_hpin_kept -= it->first;
            g_hpin_free.erase(it);
            return AMR_OK;
        }
    }
    AMR_HIP(hipHostMalloc(out, bytes, hipHostMallocPortable));
    *got = bytes;
    return AMR_OK;
}

static void pinned_put(void *p, size_t bytes)
{
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(g_hpin_mu);
        if (g_hpin_kept + bytes <= ((size_t)256 << 20)) {
            g_hpin_free.emplace(bytes, p);
            g_hpin_kept += bytes;
            return;
        }
    }
    (void)hipHostFree(p);
}

static void free_model(amr_model *m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->pool) amr_pool_detach_model(m);
    if (m->self_pool) amr_pool_free(m->self_pool);
    m->self_pool = nullptr;
    amr_coarse_free(m);
    amr_recovery_free(m);
    if (m->own_stream) (void)hipStreamSynchronize(m->own_stream);
    void *ptrs[] = {m->X,    m->conn, m->emat, m->mats,  m->dNdR, m->wip,  m->st_cur, m->st_com, m->Ke,
                    m->rowptr, m->col, m->diag, m->cptr,  m->csrc, m->fptr, m->fsrc,   m->Kv,     m->Kv32,
                    m->Kflag, m->dinv, m->dinv32, m->mask, m->eqd,  m->dF,   m->estat,  m->err,    m->part,
                    m->scal, m->binv, m->binv32, m->Me, m->Mv, m->Mflag, m->Nd, m->rho_d, m->mvx, m->mvy, m->Ksave, m->col16, m->Kpred};
    for (void *p : ptrs) amr_release(m, p);
    amr_release(m, m->v_block);
    amr_release(m, m->f_block);
    amr_release(m, m->sv_block);
    for (hipEvent_t e : m->ev)
        if (e) (void)hipEventDestroy(e);
    pinned_put(m->fin_h, sizeof(amr_model::FinHost));
    if (m->sb_event) {
        (void)hipEventSynchronize(m->sb_event);
        (void)hipEventDestroy(m->sb_event);
    }
    pinned_put(m->sb_pin, m->sb_pin_bytes);
    if (m->ev_handoff) (void)hipEventDestroy(m->ev_handoff);
    if (m->ev_batch) (void)hipEventDestroy(m->ev_batch);
    if (m->fin_event) (void)hipEventDestroy(m->fin_event);
    if (m->own_stream) (void)hipStreamDestroy(m->own_stream);
    delete m;
}

int amr_pattern_host(amr_model *m)
{
    if ((int64_t)m->pat.rowptr.size() == m->nnodes + 1) return AMR_OK;
    m->pat.rowptr.resize(m->nnodes + 1);
    m->pat.col.resize(m->pat.nblk);
    m->pat.diag.resize(m->nnodes);
    TRY(d2h(m->pat.rowptr.data(), m->rowptr, m->pat.rowptr.size(), m->stream));
    TRY(d2h(m->pat.col.data(), m->col, m->pat.col.size(), m->stream));
    TRY(d2h(m->pat.diag.data(), m->diag, m->pat.diag.size(), m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    return AMR_OK;
}

// 16-bit column offsets (column - row) for the PCG SpMV, when every block's
// offset fits; otherwise the SpMV reads the 32-bit columns.
static int build_col16(amr_model *m, hipStream_t s)
{
    TRY(amr_pattern_host(m));
    std::vector<int16_t> c16(m->pat.col.size());
    for (int64_t r = 0; r < m->nnodes; r++)
        for (int32_t k = m->pat.rowptr[r]; k < m->pat.rowptr[r + 1]; k++) {
            const int64_t d = (int64_t)m->pat.col[k] - r;
            if (d < -32768 || d > 32767) return AMR_OK;
            c16[k] = (int16_t)d;
        }
    if (c16.empty()) return AMR_OK;
    TRY(dalloc(m, &m->col16, c16.size()));
    TRY(h2d(m->col16, c16.data(), c16.size(), s));
    AMR_HIP(hipStreamSynchronize(s));   // c16 goes out of scope
    return AMR_OK;
}

static int upload_dofs(amr_model *m, const int32_t *eq_id, int64_t ndofs, int64_t nu)
{
    if (ndofs < 0 || nu < 0 || nu > ndofs) {
        amr_set_error("invalid ndofs/nu");
        return AMR_ERR_ARG;
    }
    m->ndofs = ndofs;
    m->nu = nu;
    m->eq.assign(eq_id, eq_id + m->nnodes * 3);
    m->free_h.assign(m->nnodes * 3, 0);
    std::vector<double> mask(m->nnodes * 3, 0.0);
    std::vector<char> seen(ndofs, 0);
    for (int64_t k = 0; k < m->nnodes * 3; k++) {
        int32_t q = eq_id[k];
        if (q < -1 || q >= ndofs || (q >= 0 && seen[q])) {
            amr_set_error("eq_id is not an injective map into [0, ndofs)");
            return AMR_ERR_ARG;
        }
        if (q >= 0) {
            seen[q] = 1;
            if (q < nu) {
                m->free_h[k] = 1;
                mask[k] = 1.0;
            }
        }
    }
    TRY(h2d(m->mask, mask.data(), mask.size(), m->stream));
    m->k_version++;
    m->dof_version++;
    TRY(h2d(m->eqd, eq_id, (size_t)m->nnodes * 3, m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    m->assembled = false;
    return AMR_OK;
}

// The stage driver's device vectors, pinned staging and event: allocated
// with the model (so that starting a stage allocates nothing).
static int stage_alloc(amr_model *m)
{
    const size_t n = (size_t)m->nnodes * 3;
    if (!m->sv_block) {
        // the stage vectors as slices of one block (one allocation)
        const size_t np = (n + 63) / 64 * 64, ns = sizeof(m->sv) / sizeof(m->sv[0]);
        TRY(dalloc(m, &m->sv_block, np * ns));
        for (size_t k = 0; k < ns; k++) m->sv[k] = m->sv_block + np * k;
    }
    if (!m->sb_event) {
        AMR_HIP(hipEventCreateWithFlags(&m->sb_event, hipEventDisableTiming));
        TRY(pinned_get(2 * n * sizeof(double), (void **)&m->sb_pin, &m->sb_pin_bytes));
        m->sb_fresh = true;
    }
    if (!m->fin_h) {
        size_t got = 0;
        TRY(pinned_get(sizeof(amr_model::FinHost), (void **)&m->fin_h, &got));
    }
    return AMR_OK;
}

extern "C" int amr_model_create(amr_model **out, int device, int etype, int nips, int64_t nnodes,
                                const double *coords, int64_t ne, const int32_t *conn, int nmats,
                                const int32_t *mat_kind, const double *mat_params, const int32_t *emat,
                                const int32_t *eq_id, int64_t ndofs, int64_t nu)
{
    PhaseRange range_("amr_model_create");
    if (!out || !coords || !conn || !mat_kind || !mat_params || !emat || !eq_id || nnodes <= 0 || ne < 0 ||
        nmats <= 0) {
        amr_set_error("amr_model_create: null pointer or empty model");
        return AMR_ERR_ARG;
    }
    *out = nullptr;
    const int nn = nn_of(etype);
    if (nn < 0) {
        amr_set_error("amr_model_create: unknown element type " + std::to_string(etype));
        return AMR_ERR_ARG;
    }
    double R[AMR_MAXIP * 3], w[AMR_MAXIP];
    const int nq = ip_rule(etype, nips, R, w);
    if (nq < 0) {
        amr_set_error("amr_model_create: element type " + std::to_string(etype) + " has no " +
                      std::to_string(nips) + "-point rule");
        return AMR_ERR_ARG;
    }
    for (int64_t e = 0; e < ne; e++)
        if (emat[e] < 0 || emat[e] >= nmats) {
            amr_set_error("amr_model_create: element material index out of range");
            return AMR_ERR_ARG;
        }
    std::vector<AmrMat> mats(nmats);
    bool all_el = true;
    for (int i = 0; i < nmats; i++) {
        if (mat_kind[i] < AMR_MAT_LE || mat_kind[i] > AMR_MAT_DP) {
            amr_set_error("amr_model_create: unknown material kind");
            return AMR_ERR_ARG;
        }
        mats[i].kind = mat_kind[i];
        mats[i].pad = 0;
        for (int k = 0; k < 6; k++) mats[i].p[k] = mat_params[i * 6 + k];
        if (mat_kind[i] != AMR_MAT_LE) all_el = false;
    }

    std::unique_ptr<amr_model, void (*)(amr_model *)> m(new amr_model, free_model);
    m->etype = etype;
    m->nn = nn;
    m->nips = nq;
    m->device = device;
    m->nnodes = nnodes;
    m->ne = ne;
    m->nip_total = ne * nq;
    m->nmats = nmats;
    m->all_elastic = all_el;

    AMR_HIP(hipSetDevice(device));
    AMR_HIP(hipDeviceGetAttribute(&m->ncu, hipDeviceAttributeMultiprocessorCount, device));
    AMR_HIP(hipStreamCreateWithFlags(&m->own_stream, hipStreamNonBlocking));
    m->stream = m->own_stream;
    const size_t n3 = (size_t)nnodes * 3;
    TRY(dalloc(m.get(), &m->X, n3));
    TRY(dalloc(m.get(), &m->conn, (size_t)ne * nn));
    TRY(dalloc(m.get(), &m->emat, (size_t)ne));
    TRY(dalloc(m.get(), &m->mats, (size_t)nmats));
    TRY(dalloc(m.get(), &m->dNdR, (size_t)nq * nn * 3));
    TRY(dalloc(m.get(), &m->wip, (size_t)nq));
    TRY(dalloc(m.get(), &m->st_cur, (size_t)AMR_NSTATE * m->nip_total));
    TRY(dalloc(m.get(), &m->st_com, (size_t)AMR_NSTATE * m->nip_total));
    TRY(dalloc(m.get(), &m->Ke, (size_t)ne * nn * nn * 9));
    TRY(dalloc(m.get(), &m->rowptr, (size_t)nnodes + 1));
    TRY(dalloc(m.get(), &m->diag, (size_t)nnodes));
    TRY(dalloc(m.get(), &m->fptr, (size_t)nnodes + 1));
    TRY(dalloc(m.get(), &m->fsrc, (size_t)ne * nn));
    TRY(dalloc(m.get(), &m->part, AMR_PART_DBL));
    hipStream_t s = m->stream;
    TRY(h2d(m->conn, conn, (size_t)ne * nn, s));
    // the sparsity pattern and gather maps, on the device (host builder for
    // nodes with more contributions than one workgroup sorts)
    int prc = amr_build_pattern_device(m.get());
    if (prc == AMR_PATTERN_ON_HOST) {
        TRY(amr_build_pattern(nn, nnodes, ne, conn, m->pat));
        TRY(dalloc(m.get(), &m->col, (size_t)m->pat.nblk));
        TRY(dalloc(m.get(), &m->cptr, (size_t)m->pat.nblk + 1));
        TRY(dalloc(m.get(), &m->csrc, (size_t)m->pat.ncontrib));
        TRY(h2d(m->rowptr, m->pat.rowptr.data(), m->pat.rowptr.size(), s));
        TRY(h2d(m->col, m->pat.col.data(), m->pat.col.size(), s));
        TRY(h2d(m->diag, m->pat.diag.data(), m->pat.diag.size(), s));
        TRY(h2d(m->cptr, m->pat.cptr.data(), m->pat.cptr.size(), s));
        TRY(h2d(m->csrc, m->pat.csrc.data(), m->pat.csrc.size(), s));
        TRY(h2d(m->fptr, m->pat.fptr.data(), m->pat.fptr.size(), s));
        TRY(h2d(m->fsrc, m->pat.fsrc.data(), m->pat.fsrc.size(), s));
        TRY(build_col16(m.get(), s));
        AMR_HIP(hipStreamSynchronize(s));   // the host vectors below are released
        std::vector<int32_t>().swap(m->pat.csrc);
        std::vector<int64_t>().swap(m->pat.cptr);
        std::vector<int32_t>().swap(m->pat.fsrc);
        std::vector<int32_t>().swap(m->pat.fptr);
    } else if (prc) {
        return prc;
    }
    TRY(dalloc(m.get(), &m->Kv, (size_t)kv_size(m->pat.nblk)));
    TRY(dalloc(m.get(), &m->Kv32, (size_t)kv_size(m->pat.nblk)));
    TRY(dalloc(m.get(), &m->Kflag, (size_t)kv_size(m->pat.nblk)));
    TRY(dalloc(m.get(), &m->dinv, n3));
    TRY(dalloc(m.get(), &m->dinv32, n3));
    TRY(dalloc(m.get(), &m->binv, n3 * 3));
    TRY(dalloc(m.get(), &m->binv32, n3 * 3));
    TRY(dalloc(m.get(), &m->mask, n3));
    TRY(dalloc(m.get(), &m->eqd, n3));
    TRY(dalloc(m.get(), &m->dF, (size_t)ne * nn * 3));
    TRY(dalloc(m.get(), &m->estat, (size_t)ne));
    TRY(dalloc(m.get(), &m->err, 4));
    TRY(dalloc(m.get(), &m->scal, 64));
    {
        // the work vectors as slices of one block each (one allocation)
        const size_t n3p = ((size_t)n3 + 63) / 64 * 64;
        TRY(dalloc(m.get(), &m->v_block, n3p * 12));
        TRY(dalloc(m.get(), &m->f_block, n3p * 8));
        for (int k = 0; k < 12; k++) m->v[k] = m->v_block + n3p * k;
        for (int k = 0; k < 8; k++) m->f[k] = m->f_block + n3p * k;
    }

    std::vector<double> dNdR((size_t)nq * nn * 3);
    for (int q = 0; q < nq; q++) shape_deriv(etype, R + 3 * q, dNdR.data() + (size_t)q * nn * 3);
    m->Nip.resize((size_t)nq * nn);
    for (int q = 0; q < nq; q++) shape_func(etype, R + 3 * q, m->Nip.data() + (size_t)q * nn);
    m->elem_kind = -1;
    for (int64_t e = 0; e < ne; e++) {
        const int k = mat_kind[emat[e]];
        if (e == 0) m->elem_kind = k;
        else if (k != m->elem_kind) { m->elem_kind = -2; break; }
    }
    if (m->elem_kind == -2) {
        m->ekind.resize(ne);
        for (int64_t e = 0; e < ne; e++) m->ekind[e] = (int8_t)mat_kind[emat[e]];
    }
    TRY(h2d(m->X, coords, n3, s));
    m->Xh.assign(coords, coords + n3);
    TRY(h2d(m->emat, emat, (size_t)ne, s));
    TRY(h2d(m->mats, mats.data(), mats.size(), s));
    TRY(h2d(m->dNdR, dNdR.data(), dNdR.size(), s));
    TRY(h2d(m->wip, w, (size_t)nq, s));
    AMR_HIP(hipMemsetAsync(m->st_cur, 0, sizeof(double) * AMR_NSTATE * m->nip_total, s));
    AMR_HIP(hipMemsetAsync(m->st_com, 0, sizeof(double) * AMR_NSTATE * m->nip_total, s));
    AMR_HIP(hipMemsetAsync(m->scal, 0, sizeof(double) * 64, s));
    AMR_HIP(hipMemsetAsync(m->err, 0, 4 * sizeof(int32_t), s));
    TRY(upload_dofs(m.get(), eq_id, ndofs, nu));
    TRY(stage_alloc(m.get()));
    *out = m.release();
    return AMR_OK;
}

extern "C" int amr_model_destroy(amr_model *m)
{
    free_model(m);
    return AMR_OK;
}

extern "C" int amr_model_set_dofs(amr_model *m, const int32_t *eq_id, int64_t ndofs, int64_t nu)
{
    if (!m || !eq_id) return AMR_ERR_ARG;
    if (amr_pool_is_member(m)) {
        amr_set_error("the model is solving in a pool; its matrix and solver vectors are in use");
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipSetDevice(m->device));
    return upload_dofs(m, eq_id, ndofs, nu);
}

extern "C" int amr_model_info(amr_model *m, int64_t *out)
{
    if (!m || !out) return AMR_ERR_ARG;
    out[0] = m->nnodes;
    out[1] = m->ne;
    out[2] = m->nips;
    out[3] = m->pat.nblk;
    out[4] = m->pat.nblk * 9;
    out[5] = m->pat.ncontrib;
    out[6] = m->ndofs;
    out[7] = m->nu;
    return AMR_OK;
}

// ----------------------------------------------------------------- state

extern "C" int amr_state_set(amr_model *m, const double *st)
{
    if (!m || !st) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    const int64_t n = m->nip_total;
    std::vector<double> soa((size_t)n * AMR_NSTATE);
    for (int64_t i = 0; i < n; i++)
        for (int f = 0; f < AMR_NSTATE; f++) soa[f * n + i] = st[i * AMR_NSTATE + f];
    TRY(h2d(m->st_cur, soa.data(), soa.size(), m->stream));
    TRY(h2d(m->st_com, soa.data(), soa.size(), m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    return AMR_OK;
}

// IP states live field-major on the device (field f of IP i at f * n + i);
// the host layout is IP-major (i * AMR_NSTATE + f).  The read-back
// transposes on the device into a staging buffer, one per device, grown on
// demand and kept until the process ends, and copies it out in one piece.
__global__ void state_aos_kernel(int64_t n, const double *__restrict__ soa, double *__restrict__ aos)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n * AMR_NSTATE) return;
    const int64_t i = k / AMR_NSTATE;
    const int f = (int)(k - i * AMR_NSTATE);
    aos[k] = soa[(int64_t)f * n + i];
}

static std::mutex g_stage_mu;
static double *g_stage[64] = {};
static size_t g_stage_cap[64] = {};

// Read-backs of result vectors go through pinned host memory (one buffer
// per device, grown as needed; a copy into pageable memory takes the
// runtime's slower path).  The caller holds g_pin_mu until it has copied
// the data out.
static std::mutex g_pin_mu;
static double *g_pin[64] = {};
static size_t g_pin_cap[64] = {};

static int pinned_buffer(int device, size_t count, double **out)
{
    if (g_pin_cap[device] < count) {
        if (g_pin[device]) AMR_HIP(hipHostFree(g_pin[device]));
        g_pin[device] = nullptr;
        g_pin_cap[device] = 0;
        const size_t cap = std::max(count, (size_t)1 << 16);
        AMR_HIP(hipHostMalloc((void **)&g_pin[device], cap * sizeof(double), hipHostMallocDefault));
        g_pin_cap[device] = cap;
    }
    *out = g_pin[device];
    return AMR_OK;
}

extern "C" int amr_state_get(amr_model *m, double *st, int committed)
{
    if (!m || !st) return AMR_ERR_ARG;
    if (m->device < 0 || m->device >= 64) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    const int64_t n = m->nip_total;
    const size_t cnt = (size_t)n * AMR_NSTATE;
    if (cnt == 0) return AMR_OK;
    // the staging buffer is shared by the device's models: one read-back at a time
    std::lock_guard<std::mutex> lk(g_stage_mu);
    double *&buf = g_stage[m->device];
    if (g_stage_cap[m->device] < cnt) {
        if (buf) AMR_HIP(hipFree(buf));
        buf = nullptr;
        g_stage_cap[m->device] = 0;
        AMR_HIP(hipMalloc((void **)&buf, cnt * sizeof(double)));
        g_stage_cap[m->device] = cnt;
    }
    const unsigned g = (unsigned)((cnt + 255) / 256);
    hipLaunchKernelGGL(state_aos_kernel, dim3(g), dim3(256), 0, m->stream, n, committed ? m->st_com : m->st_cur,
                       buf);
    AMR_HIP(hipGetLastError());
    std::lock_guard<std::mutex> lp(g_pin_mu);
    double *pin;
    TRY(pinned_buffer(m->device, cnt, &pin));
    TRY(d2h(pin, buf, cnt, m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    std::memcpy(st, pin, cnt * sizeof(double));
    return AMR_OK;
}

extern "C" int amr_state_commit(amr_model *m)
{
    if (!m) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    AMR_HIP(hipMemcpyAsync(m->st_com, m->st_cur, sizeof(double) * AMR_NSTATE * m->nip_total,
                           hipMemcpyDeviceToDevice, m->stream));
    return AMR_OK;
}

extern "C" int amr_state_restore(amr_model *m)
{
    if (!m) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    AMR_HIP(hipMemcpyAsync(m->st_cur, m->st_com, sizeof(double) * AMR_NSTATE * m->nip_total,
                           hipMemcpyDeviceToDevice, m->stream));
    return AMR_OK;
}

// -------------------------------------------------------------- mount_K

extern "C" int amr_assemble(amr_model *m)
{
    PhaseRange range_("mount_K");
    if (!m) return AMR_ERR_ARG;
    if (amr_pool_is_member(m)) {
        amr_set_error("the model is solving in a pool; its matrix and solver vectors are in use");
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipSetDevice(m->device));
    AMR_HIP(hipMemsetAsync(m->err, 0x7f, sizeof(int32_t), m->stream));  // "no bad element"
    TRY(amr_launch_ke_path(m, AMR_KE_EXACT));   // mount_K's pattern comes from the exact element matrices
    TRY(amr_launch_assemble(m, true));
    int32_t err = 0;
    TRY(d2h(&err, m->err, 1, m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    if (err != 0x7f7f7f7f) {
        amr_set_error("Negative Jacobian determinant in cell " + std::to_string(err));
        return AMR_ERR_JACOBIAN;
    }
    m->assembled = true;
    return AMR_OK;
}

extern "C" int amr_get_Ke(amr_model *m, double *Ke)
{
    if (!m || !Ke) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    const int nn = m->nn, nd = 3 * nn;
    std::vector<double> blk((size_t)m->ne * nn * nn * 9);
    TRY(d2h(blk.data(), m->Ke, blk.size(), m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    for (int64_t e = 0; e < m->ne; e++)
        for (int a = 0; a < nn; a++)
            for (int b = 0; b < nn; b++)
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++)
                        Ke[(e * nd + 3 * a + r) * nd + 3 * b + c] = blk[((e * nn + a) * nn + b) * 9 + r * 3 + c];
    return AMR_OK;
}

// CSC over the DOF numbering of a matrix held in the node-block layout
// (values and survivor flags in the kv_at layout).
static int bsr_to_csc(amr_model *m, const double *dv, const uint8_t *dflag, int64_t *nnz, int64_t *colptr,
                      int32_t *rowval, double *nzval)
{
    AMR_HIP(hipSetDevice(m->device));
    TRY(amr_pattern_host(m));
    const int64_t nb = m->pat.nblk;
    std::vector<double> val((size_t)kv_size(nb));
    std::vector<uint8_t> fl((size_t)kv_size(nb));
    TRY(d2h(val.data(), dv, val.size(), m->stream));
    TRY(d2h(fl.data(), dflag, fl.size(), m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    std::vector<int64_t> cnt(m->ndofs + 1, 0);
    int64_t total = 0;
    for (int64_t A = 0; A < m->nnodes; A++)
        for (int32_t k = m->pat.rowptr[A]; k < m->pat.rowptr[A + 1]; k++) {
            const int64_t B = m->pat.col[k];
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    if (!fl[kv_at(k, r * 3 + c)]) continue;
                    const int32_t i = m->eq[A * 3 + r], j = m->eq[B * 3 + c];
                    if (i < 0 || j < 0) continue;
                    cnt[j + 1]++;
                    total++;
                }
        }
    *nnz = total;
    if (!colptr) return AMR_OK;
    if (!rowval || !nzval) return AMR_ERR_ARG;
    for (int64_t j = 0; j < m->ndofs; j++) cnt[j + 1] += cnt[j];
    std::copy(cnt.begin(), cnt.end(), colptr);
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t A = 0; A < m->nnodes; A++)
        for (int32_t k = m->pat.rowptr[A]; k < m->pat.rowptr[A + 1]; k++) {
            const int64_t B = m->pat.col[k];
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    if (!fl[kv_at(k, r * 3 + c)]) continue;
                    const int32_t i = m->eq[A * 3 + r], j = m->eq[B * 3 + c];
                    if (i < 0 || j < 0) continue;
                    int64_t p = pos[j]++;
                    rowval[p] = i;
                    nzval[p] = val[kv_at(k, r * 3 + c)];
                }
        }
    // sort row ids within each column
    std::vector<std::pair<int32_t, double>> tmp;
    for (int64_t j = 0; j < m->ndofs; j++) {
        const int64_t b = colptr[j], e = colptr[j + 1];
        tmp.resize(e - b);
        for (int64_t p = b; p < e; p++) tmp[p - b] = {rowval[p], nzval[p]};
        std::sort(tmp.begin(), tmp.end(),
                  [](const std::pair<int32_t, double> &x, const std::pair<int32_t, double> &y) {
                      return x.first < y.first;
                  });
        for (int64_t p = b; p < e; p++) {
            rowval[p] = tmp[p - b].first;
            nzval[p] = tmp[p - b].second;
        }
    }
    return AMR_OK;
}

extern "C" int amr_get_K_csc(amr_model *m, int64_t *nnz, int64_t *colptr, int32_t *rowval, double *nzval)
{
    if (!m || !nnz) return AMR_ERR_ARG;
    if (!m->assembled) {
        amr_set_error("amr_get_K_csc: K has not been assembled");
        return AMR_ERR_ARG;
    }
    return bsr_to_csc(m, m->Kv, m->Kflag, nnz, colptr, rowval, nzval);
}

extern "C" int amr_assemble_M(amr_model *m, const double *rho)
{
    if (!m || !rho) return AMR_ERR_ARG;
    for (int64_t e = 0; e < m->ne; e++)
        if (!(rho[e] >= 0.0)) {
            amr_set_error("amr_assemble_M: rho must be >= 0 (element " + std::to_string(e + 1) + ")");
            return AMR_ERR_ARG;
        }
    AMR_HIP(hipSetDevice(m->device));
    const size_t nq = (size_t)m->nips, nn = (size_t)m->nn;
    if (!m->Me) TRY(dalloc(m, &m->Me, (size_t)m->ne * nn * nn));
    if (!m->Mv) TRY(dalloc(m, &m->Mv, (size_t)kv_size(m->pat.nblk)));
    if (!m->Mflag) TRY(dalloc(m, &m->Mflag, (size_t)kv_size(m->pat.nblk)));
    if (!m->Nd) TRY(dalloc(m, &m->Nd, nq * nn));
    if (!m->rho_d) TRY(dalloc(m, &m->rho_d, (size_t)m->ne));
    TRY(h2d(m->Nd, m->Nip.data(), m->Nip.size(), m->stream));
    TRY(h2d(m->rho_d, rho, (size_t)m->ne, m->stream));
    AMR_HIP(hipMemsetAsync(m->err, 0x7f, sizeof(int32_t), m->stream));  // "no bad element"
    TRY(amr_launch_mass(m));
    int32_t err = 0;
    TRY(d2h(&err, m->err, 1, m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    if (err != 0x7f7f7f7f) {
        amr_set_error("Negative Jacobian determinant in cell " + std::to_string(err));
        return AMR_ERR_JACOBIAN;
    }
    m->m_assembled = true;
    return AMR_OK;
}

extern "C" int amr_spmv_col_bytes(amr_model *m, int *bytes)
{
    if (!m || !bytes) return AMR_ERR_ARG;
    *bytes = m->col16 ? 2 : 4;
    return AMR_OK;
}

extern "C" int amr_set_spmv_col_bytes(amr_model *m, int bytes)
{
    if (!m || (bytes != 2 && bytes != 4)) {
        amr_set_error("amr_set_spmv_col_bytes: bytes must be 2 or 4");
        return AMR_ERR_ARG;
    }
    if (amr_pool_is_member(m)) {
        amr_set_error("the model is solving in a pool; its matrix is in use");
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipSetDevice(m->device));
    if (bytes == 4 && m->col16) {
        AMR_HIP(hipStreamSynchronize(m->stream));
        amr_release(m, m->col16);
        m->col16 = nullptr;
    } else if (bytes == 2 && !m->col16) {
        TRY(build_col16(m, m->stream));
        if (!m->col16) {
            amr_set_error("amr_set_spmv_col_bytes: a block's column offset does not fit 16 bits");
            return AMR_ERR_ARG;
        }
    }
    return AMR_OK;
}

extern "C" int amr_combine_KM(amr_model *m, double a, double b)
{
    if (!m) return AMR_ERR_ARG;
    if (!m->assembled || !m->m_assembled) {
        amr_set_error("amr_combine_KM: assemble K (amr_assemble) and M (amr_assemble_M) first");
        return AMR_ERR_ARG;
    }
    if (amr_pool_is_member(m)) {
        amr_set_error("the model is solving in a pool; its matrix and solver vectors are in use");
        return AMR_ERR_ARG;
    }
    if (!std::isfinite(a) || !std::isfinite(b)) {
        amr_set_error("amr_combine_KM: non-finite coefficient");
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipSetDevice(m->device));
    // K as assembled stays available to amr_matvec (which 2): the Newmark
    // iteration multiplies the damping matrix alpha M + beta K after the solve
    if (!m->Ksave) TRY(dalloc(m, &m->Ksave, (size_t)kv_size(m->pat.nblk)));
    AMR_HIP(hipMemcpyAsync(m->Ksave, m->Kv, sizeof(double) * kv_size(m->pat.nblk), hipMemcpyDeviceToDevice,
                           m->stream));
    m->ksave_version = m->k_version;
    TRY(amr_launch_combine_km(m, a, b));
    AMR_HIP(hipStreamSynchronize(m->stream));
    return AMR_OK;
}

extern "C" int amr_matvec(amr_model *m, int which, const double *x, double *y)
{
    if (!m || !x || !y || which < 0 || which > 2) return AMR_ERR_ARG;
    if (which == 0 ? !m->assembled : which == 1 ? !m->m_assembled : m->ksave_version < 0) {
        amr_set_error(which == 0   ? "amr_matvec: K has not been assembled"
                      : which == 1 ? "amr_matvec: M has not been assembled"
                                   : "amr_matvec: no K was combined with M");
        return AMR_ERR_ARG;
    }
    if (amr_pool_is_member(m)) {
        amr_set_error("the model is solving in a pool; its matrix and solver vectors are in use");
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipSetDevice(m->device));
    const size_t n3 = (size_t)m->nnodes * 3;
    if (!m->mvx) TRY(dalloc(m, &m->mvx, n3));
    if (!m->mvy) TRY(dalloc(m, &m->mvy, n3));
    // eq order -> node order (DOFs without an eq id hold 0)
    std::vector<double> xn(n3, 0.0), yn(n3);
    for (size_t i = 0; i < n3; i++)
        if (m->eq[i] >= 0) xn[i] = x[m->eq[i]];
    TRY(h2d(m->mvx, xn.data(), n3, m->stream));
    TRY(amr_spmv_vals(m, which == 0 ? m->Kv : which == 1 ? m->Mv : m->Ksave, m->mvx, m->mvy, 0));
    TRY(d2h(yn.data(), m->mvy, n3, m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    for (size_t i = 0; i < n3; i++)
        if (m->eq[i] >= 0) y[m->eq[i]] = yn[i];
    return AMR_OK;
}

extern "C" int amr_get_M_csc(amr_model *m, int64_t *nnz, int64_t *colptr, int32_t *rowval, double *nzval)
{
    if (!m || !nnz) return AMR_ERR_ARG;
    if (!m->m_assembled) {
        amr_set_error("amr_get_M_csc: M has not been assembled");
        return AMR_ERR_ARG;
    }
    return bsr_to_csc(m, m->Mv, m->Mflag, nnz, colptr, rowval, nzval);
}

// ------------------------------------------------------- solve_system!

static void eq_to_node(const amr_model *m, const double *eqv, double *node)
{
    for (int64_t k = 0; k < m->nnodes * 3; k++) node[k] = m->eq[k] >= 0 ? eqv[m->eq[k]] : 0.0;
}

static void eq_to_node(const amr_model *m, const double *eqv, std::vector<double> &node)
{
    node.resize(m->nnodes * 3);
    eq_to_node(m, eqv, node.data());
}

static void node_to_eq(const amr_model *m, const double *node, double *eqv)
{
    for (int64_t k = 0; k < m->nnodes * 3; k++)
        if (m->eq[k] >= 0) eqv[m->eq[k]] = node[k];
}

__global__ static void prep_rhs_kernel(int64_t n, const double *__restrict__ mask, const double *__restrict__ u,
                                       double *__restrict__ u2)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u2[i] = mask[i] != 0.0 ? 0.0 : u[i];
}

__global__ static void rhs_kernel(int64_t n, const double *__restrict__ mask, const double *__restrict__ f,
                                  const double *__restrict__ ky, double *__restrict__ b)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = mask[i] != 0.0 ? f[i] - ky[i] : 0.0;
}

__global__ static void combine_kernel(int64_t n, const double *__restrict__ mask, const double *__restrict__ x,
                                      double *__restrict__ u2)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && mask[i] != 0.0) u2[i] = x[i];
}

static amr_solve_opts default_opts()
{
    amr_solve_opts o;
    o.precision = AMR_PREC_FP64;
    o.precond = AMR_PRECOND_JACOBI;
    o.maxiter = 200000;
    o.check_every = 32;
    o.rtol = 1e-12;
    o.inner_rtol = 1e-4;
    o.replace_ratio = 0.0;
    o.drift_limit = 0.0;
    return o;
}

extern "C" int amr_solve_system(amr_model *m, double *U, double *F, const amr_solve_opts *opts,
                                amr_solve_info *info)
{
    PhaseRange range_("solve_system!");
    if (!m || !U || !F) return AMR_ERR_ARG;
    if (amr_pool_is_member(m)) {
        amr_set_error("the model is solving in a pool; its matrix and solver vectors are in use");
        return AMR_ERR_ARG;
    }
    if (!m->assembled) {
        amr_set_error("amr_solve_system: K has not been assembled");
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipSetDevice(m->device));
    amr_solve_opts o = opts ? *opts : default_opts();
    const int64_t n = m->nnodes * 3;
    std::vector<double> un, fn;
    eq_to_node(m, U, un);
    eq_to_node(m, F, fn);
    double *du = m->v[6], *df = m->v[7], *u2 = m->v[8], *ky = m->v[9], *b = m->v[10], *x = m->v[11];
    hipStream_t s = m->stream;
    TRY(h2d(du, un.data(), (size_t)n, s));
    TRY(h2d(df, fn.data(), (size_t)n, s));
    const unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(prep_rhs_kernel, dim3(g), dim3(256), 0, s, n, m->mask, du, u2);
    TRY(amr_spmv_full(m, u2, ky, 0));
    hipLaunchKernelGGL(rhs_kernel, dim3(g), dim3(256), 0, s, n, m->mask, df, ky, b);
    AMR_HIP(hipGetLastError());
    amr_solve_info inf{};
    int rc = AMR_OK;
    if (m->nu > 0) rc = amr_pcg_solve(m, b, x, &o, &inf);
    if (info) *info = inf;
    if (rc) return rc;
    if (m->nu > 0) hipLaunchKernelGGL(combine_kernel, dim3(g), dim3(256), 0, s, n, m->mask, x, u2);
    TRY(amr_spmv_full(m, u2, ky, 0));
    std::vector<double> uo(n), ko(n);
    TRY(d2h(uo.data(), u2, (size_t)n, s));
    TRY(d2h(ko.data(), ky, (size_t)n, s));
    AMR_HIP(hipStreamSynchronize(s));
    double umax = 0.0;
    for (int64_t k = 0; k < n; k++)
        if (m->free_h[k]) {
            if (!std::isfinite(uo[k])) umax = INFINITY;
            umax = std::max(umax, std::fabs(uo[k]));
        }
    if (umax > 1e8) {
        amr_set_error("\nsolve_system!: Possible syngular matrix");
        return AMR_ERR_SOLVE;
    }
    for (int64_t k = 0; k < n; k++) {
        const int32_t q = m->eq[k];
        if (q < 0) continue;
        if (m->free_h[k]) U[q] = uo[k];
        else F[q] = ko[k];
    }
    return AMR_OK;
}

// -------------------------------------------------------- update_state!

extern "C" int amr_update_state(amr_model *m, const double *dUt, double *dFin)
{
    PhaseRange range_("update_state!");
    if (!m || !dUt || !dFin) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    const int64_t n = m->nnodes * 3;
    std::vector<double> un;
    eq_to_node(m, dUt, un);
    double *du = m->v[6], *fo = m->v[7];
    TRY(h2d(du, un.data(), (size_t)n, m->stream));
    TRY(amr_launch_update(m, du, fo));
    std::vector<double> fn(n);
    std::vector<int32_t> st(m->ne);
    TRY(d2h(fn.data(), fo, (size_t)n, m->stream));
    TRY(d2h(st.data(), m->estat, (size_t)m->ne, m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    std::fill(dFin, dFin + m->ndofs, 0.0);
    node_to_eq(m, fn.data(), dFin);
    for (int64_t e = 0; e < m->ne; e++)
        if (st[e]) {
            if (st[e] == 2) {
                amr_set_error("Negative Jacobian determinant in cell " + std::to_string(e + 1));
                return AMR_ERR_JACOBIAN;
            }
            amr_set_error("VonMisses: Negative value for √J2D (cell " + std::to_string(e + 1) + ")");
            return AMR_ERR_MATERIAL;
        }
    for (int64_t i = 0; i < m->ndofs; i++)
        if (std::isnan(dFin[i])) {
            amr_set_error("solve_system!: NaN values in internal forces vector");
            return AMR_ERR_NAN;
        }
    return AMR_OK;
}

// ------------------------------------------------ device-resident stage

extern "C" int amr_stage_begin(amr_model *m, const double *Uex, const double *Fex)
{
    if (!m || !Uex || !Fex) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    const size_t n = (size_t)m->nnodes * 3;
    TRY(stage_alloc(m));
    // the pinned host buffer lives in the model until the copies from it are
    // done (waited for here before the next stage overwrites it)
    if (!m->sb_fresh) AMR_HIP(hipEventSynchronize(m->sb_event));
    m->sb_fresh = false;
    eq_to_node(m, Uex, m->sb_pin);
    eq_to_node(m, Fex, m->sb_pin + n);
    TRY(h2d(m->sv[SV_UEX], m->sb_pin, n, m->stream));
    TRY(h2d(m->sv[SV_FEX], m->sb_pin + n, n, m->stream));
    AMR_HIP(hipEventRecord(m->sb_event, m->stream));
    // SV_U .. SV_DFIN are consecutive slices: one fill
    static_assert(SV_DFIN - SV_U == 7 && SV_F == SV_U + 1 && SV_DFEX == SV_U + 2 && SV_R == SV_U + 3 &&
                      SV_RC == SV_U + 4 && SV_DUA == SV_U + 5 && SV_DUI == SV_U + 6,
                  "stage vectors zeroed at a stage's start");
    TRY(amr_vec_fill_n(m, m->sv[SV_U], (int64_t)(m->sv[SV_DFIN] - m->sv[SV_U]) + (int64_t)n, 0.0));
    return AMR_OK;
}

extern "C" int amr_increment_begin(amr_model *m, double dT, double acr, int add_residual)
{
    if (!m || !m->sv[SV_U]) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    m->acr = acr;
    m->inc_version++;
    return amr_vec_inc_begin(m, dT, acr, add_residual);
}

extern "C" int amr_nr_prepare(amr_model *m, int it)
{
    PhaseRange range_("mount_K (NR)");
    if (!m || !m->sv[SV_U] || it < 1) return AMR_ERR_ARG;
    if (amr_pool_is_member(m)) {
        amr_set_error("the model is solving in a pool; its matrix and solver vectors are in use");
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipSetDevice(m->device));
    double **v = m->sv;
    if (it > 1) TRY(amr_vec_fill(m, v[SV_DUI], 0.0));
    // mount_K
    AMR_HIP(hipMemsetAsync(m->err, 0x7f, 3 * sizeof(int32_t), m->stream));
    TRY(amr_launch_ke(m));
    TRY(amr_launch_assemble(m, false));
    m->assembled = true;
    // right-hand side of solve_system!(K, ΔUitr, Rtr, nu): R - K12 ΔUi2
    if (m->nu > 0) TRY(amr_vec_solve_prep(m, v[SV_DUI], v[SV_R], m->v[8], m->v[9], m->v[10]));
    return AMR_OK;
}

extern "C" int amr_nr_finish_begin(amr_model *m, const amr_solve_info *si)
{
    PhaseRange range_("update_state! (NR)");
    if (!m || !m->sv[SV_U]) return AMR_ERR_ARG;
    if (amr_pool_is_member(m)) {
        amr_set_error("the model is solving in a pool; its matrix and solver vectors are in use");
        return AMR_ERR_ARG;
    }
    if (m->fin_pending) {
        amr_set_error("amr_nr_finish_begin: the previous finish has not been completed");
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipSetDevice(m->device));
    TRY(stage_alloc(m));
    amr_nr_info inf{};
    double **v = m->sv;
    int rc = AMR_OK;
    if (m->nu > 0) {
        if (!si) {
            amr_set_error("amr_nr_finish: the model has unknowns; pass the outcome of its linear solve");
            return AMR_ERR_ARG;
        }
        inf.pcg_iterations = si->iterations;
        inf.relres = si->relres;
        inf.replacements = si->refinements;
        inf.fallback = si->fallback;
        if (!si->converged) rc = AMR_ERR_SOLVE;
        else TRY(amr_vec_combine(m, v[SV_X], v[SV_DUI]));
    }
    if (rc == AMR_OK) {
        // update_state!(ΔUa + ΔUitr) from the committed state
        TRY(amr_vec_add(m, v[SV_DUA], v[SV_DUI], v[SV_DUT]));
        TRY(amr_launch_update(m, v[SV_DUT], v[SV_DFIN]));
        TRY(amr_vec_nr_post_async(m, m->fin_h->s, m->fin_h->c));
        AMR_HIP(hipMemcpyAsync(m->fin_h->err, m->err, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, m->stream));
        // waited for by amr_nr_finish_end: whatever is queued on the stream
        // after this point (another pool step) is not waited for
        if (!m->fin_event) AMR_HIP(hipEventCreateWithFlags(&m->fin_event, hipEventDisableTiming));
        AMR_HIP(hipEventRecord(m->fin_event, m->stream));
    }
    m->fin_info = inf;
    m->fin_rc = rc;
    m->fin_pending = true;
    return AMR_OK;
}

// ---------------------------------------------- several models at once
//
// amr_nr_prepare_many / amr_nr_finish_begin_many: the same as calling the
// single-model entry point on each model in turn, with the element kernels
// of all the models in one launch per kernel family.  The batched launches
// run on the models' common stream (a pool's, for members that left it
// together), else on the device's batch stream, ordered after and before the
// models' own streams with events.

namespace {

// Staging of the job tables: one buffer per stream, so that an upload is
// queued behind the kernels that read that buffer's previous contents (a
// buffer shared by two streams could be overwritten while the other
// stream's kernels still read it).
struct BatchCtx {
    std::mutex mu;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::map<hipStream_t, AmrStaging> stg;
};
BatchCtx g_batch[64];

int batch_check(amr_model *const *ms, int n, int *device)
{
    for (int i = 0; i < n; i++) {
        amr_model *m = ms[i];
        if (!m || !m->sv[SV_U]) return AMR_ERR_ARG;
        if (amr_pool_is_member(m)) {
            amr_set_error("the model is solving in a pool; its matrix and solver vectors are in use");
            return AMR_ERR_ARG;
        }
        if (i > 0 && m->device != ms[0]->device) {
            amr_set_error("the models of one batched call must be on one device");
            return AMR_ERR_ARG;
        }
        for (int k = 0; k < i; k++)
            if (ms[k] == m) {
                amr_set_error("a model appears twice in one batched call");
                return AMR_ERR_ARG;
            }
    }
    *device = ms[0]->device;
    return AMR_OK;
}

// the stream of the batched launches, ordered after the models' queued work
int batch_enter(amr_model *const *ms, int n, BatchCtx &b, hipStream_t *s)
{
    bool same = true;
    for (int i = 1; i < n; i++) same = same && ms[i]->stream == ms[0]->stream;
    if (same) {
        *s = ms[0]->stream;
        return AMR_OK;
    }
    if (!b.stream) {
        AMR_HIP(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
        AMR_HIP(hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
    }
    for (int i = 0; i < n; i++) {
        amr_model *m = ms[i];
        if (!m->ev_batch) AMR_HIP(hipEventCreateWithFlags(&m->ev_batch, hipEventDisableTiming));
        AMR_HIP(hipEventRecord(m->ev_batch, m->stream));
        AMR_HIP(hipStreamWaitEvent(b.stream, m->ev_batch, 0));
    }
    *s = b.stream;
    return AMR_OK;
}

// the models' streams ordered after the batched launches
int batch_leave(amr_model *const *ms, int n, BatchCtx &b, hipStream_t s)
{
    if (s != b.stream) return AMR_OK;
    AMR_HIP(hipEventRecord(b.done, s));
    for (int i = 0; i < n; i++) AMR_HIP(hipStreamWaitEvent(ms[i]->stream, b.done, 0));
    return AMR_OK;
}

}  // namespace

extern "C" int amr_nr_prepare_many(amr_model **ms, const int *its, int n)
{
    PhaseRange range_("mount_K (NR, batched)");
    if (n <= 0) return AMR_OK;
    if (!ms || !its) return AMR_ERR_ARG;
    int dev = 0;
    TRY(batch_check(ms, n, &dev));
    for (int i = 0; i < n; i++)
        if (its[i] < 1) return AMR_ERR_ARG;
    if (dev < 0 || dev >= 64) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(dev));
    for (int i = 0; i < n; i++)
        if (its[i] > 1) TRY(amr_vec_fill(ms[i], ms[i]->sv[SV_DUI], 0.0));
    {
        BatchCtx &b = g_batch[dev];
        std::lock_guard<std::mutex> lk(b.mu);
        hipStream_t s;
        TRY(batch_enter(ms, n, b, &s));
        TRY(amr_batch_mount_K(ms, n, s, b.stg[s]));
        TRY(batch_leave(ms, n, b, s));
    }
    for (int i = 0; i < n; i++) {
        amr_model *m = ms[i];
        m->assembled = true;
        // right-hand side of solve_system!(K, ΔUitr, Rtr, nu): R - K12 ΔUi2
        if (m->nu > 0) TRY(amr_vec_solve_prep(m, m->sv[SV_DUI], m->sv[SV_R], m->v[8], m->v[9], m->v[10]));
    }
    return AMR_OK;
}

extern "C" int amr_nr_finish_begin_many(amr_model **ms, const amr_solve_info *si, int n)
{
    PhaseRange range_("update_state! (NR, batched)");
    if (n <= 0) return AMR_OK;
    if (!ms) return AMR_ERR_ARG;
    int dev = 0;
    TRY(batch_check(ms, n, &dev));
    if (dev < 0 || dev >= 64) return AMR_ERR_ARG;
    for (int i = 0; i < n; i++) {
        if (ms[i]->fin_pending) {
            amr_set_error("amr_nr_finish_begin: the previous finish has not been completed");
            return AMR_ERR_ARG;
        }
        if (ms[i]->nu > 0 && !si) {
            amr_set_error("amr_nr_finish: the model has unknowns; pass the outcome of its linear solve");
            return AMR_ERR_ARG;
        }
    }
    AMR_HIP(hipSetDevice(dev));
    std::vector<amr_model *> ok;
    std::vector<const double *> dU;
    std::vector<double *> dFin;
    std::vector<amr_nr_info> inf(n);
    std::vector<int> rcs(n, AMR_OK);
    for (int i = 0; i < n; i++) {
        amr_model *m = ms[i];
        TRY(stage_alloc(m));
        double **v = m->sv;
        if (m->nu > 0) {
            inf[i].pcg_iterations = si[i].iterations;
            inf[i].relres = si[i].relres;
            inf[i].replacements = si[i].refinements;
            inf[i].fallback = si[i].fallback;
            if (!si[i].converged) rcs[i] = AMR_ERR_SOLVE;
            else TRY(amr_vec_combine(m, v[SV_X], v[SV_DUI]));
        }
        if (rcs[i] == AMR_OK) {
            // update_state!(ΔUa + ΔUitr) from the committed state
            TRY(amr_vec_add(m, v[SV_DUA], v[SV_DUI], v[SV_DUT]));
            ok.push_back(m);
            dU.push_back(v[SV_DUT]);
            dFin.push_back(v[SV_DFIN]);
        }
    }
    if (!ok.empty()) {
        BatchCtx &b = g_batch[dev];
        std::lock_guard<std::mutex> lk(b.mu);
        hipStream_t s;
        TRY(batch_enter(ok.data(), (int)ok.size(), b, &s));
        TRY(amr_batch_update(ok.data(), dU.data(), dFin.data(), (int)ok.size(), s, b.stg[s]));
        TRY(batch_leave(ok.data(), (int)ok.size(), b, s));
    }
    for (amr_model *m : ok) {
        TRY(amr_vec_nr_post_async(m, m->fin_h->s, m->fin_h->c));
        AMR_HIP(hipMemcpyAsync(m->fin_h->err, m->err, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, m->stream));
        if (!m->fin_event) AMR_HIP(hipEventCreateWithFlags(&m->fin_event, hipEventDisableTiming));
        AMR_HIP(hipEventRecord(m->fin_event, m->stream));
    }
    for (int i = 0; i < n; i++) {
        ms[i]->fin_info = inf[i];
        ms[i]->fin_rc = rcs[i];
        ms[i]->fin_pending = true;
    }
    return AMR_OK;
}

extern "C" int amr_nr_finish_end(amr_model *m, amr_nr_info *info)
{
    if (!m || !m->fin_pending) return AMR_ERR_ARG;
    amr_nr_info inf = m->fin_info;
    int rc = m->fin_rc;
    m->fin_pending = false;
    if (rc == AMR_ERR_SOLVE) {
        char buf[200];
        snprintf(buf, sizeof buf, "PCG did not converge (relres %.3e after %d iterations, %d replacements)",
                 inf.relres, inf.pcg_iterations, inf.replacements);
        amr_set_error(buf);
    } else if (rc == AMR_OK) {
        AMR_HIP(hipSetDevice(m->device));
        AMR_HIP(hipEventSynchronize(m->fin_event));
        int i_res, i_errn, i_errd, i_nan;
        amr_nr_post_slots(&i_res, &i_errn, &i_errd, &i_nan);
        const double umax = m->fin_h->s[i_errn], uamax = m->fin_h->s[i_errd];
        const int32_t *herr = m->fin_h->err;
        inf.res = m->fin_h->s[i_res];
        inf.err = umax / uamax;
        if (herr[0] != 0x7f7f7f7f) {
            amr_set_error("Negative Jacobian determinant in cell " + std::to_string(herr[0]));
            rc = AMR_ERR_JACOBIAN;
        } else if (herr[2] != 0x7f7f7f7f) {
            amr_set_error("Negative Jacobian determinant in cell " + std::to_string(herr[2]));
            rc = AMR_ERR_JACOBIAN;
        } else if (herr[1] != 0x7f7f7f7f) {
            amr_set_error("VonMisses: Negative value for √J2D (cell " + std::to_string(herr[1]) + ")");
            rc = AMR_ERR_MATERIAL;
        } else if (m->fin_h->c[i_nan]) {
            amr_set_error("solve_system!: NaN values in internal forces vector");
            rc = AMR_ERR_NAN;
        } else if (umax > 1e8) {
            // max|U1| > 1e8 (src/solver.jl:68-71); ΔUi's prescribed part is the
            // essential increment, far below this bound in any sane model
            amr_set_error("\nsolve_system!: Possible syngular matrix");
            rc = AMR_ERR_SOLVE;
        }
    }
    inf.status = rc;
    if (info) *info = inf;
    return rc;
}

extern "C" int amr_nr_finish(amr_model *m, const amr_solve_info *si, amr_nr_info *info)
{
    TRY(amr_nr_finish_begin(m, si));
    return amr_nr_finish_end(m, info);
}

extern "C" int amr_set_scheme(amr_model *m, int scheme)
{
    if (!m || scheme < AMR_SCHEME_FE || scheme > AMR_SCHEME_RALSTON) {
        amr_set_error("amr_set_scheme: unknown scheme");
        return AMR_ERR_ARG;
    }
    if (amr_pool_is_member(m)) {
        amr_set_error("the model is solving in a pool; its matrix and solver vectors are in use");
        return AMR_ERR_ARG;
    }
    m->scheme = scheme;
    return AMR_OK;
}

// The error words and the NaN flag of an update_state! queued on the
// model's stream (waits for it); AMR_OK or the failure, message set.
static int update_status(amr_model *m)
{
    int32_t herr[4];
    TRY(d2h(herr, m->err, 4, m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    if (herr[0] != 0x7f7f7f7f || herr[2] != 0x7f7f7f7f) {
        amr_set_error("Negative Jacobian determinant in cell " +
                      std::to_string(herr[0] != 0x7f7f7f7f ? herr[0] : herr[2]));
        return AMR_ERR_JACOBIAN;
    }
    if (herr[1] != 0x7f7f7f7f) {
        amr_set_error("VonMisses: Negative value for √J2D (cell " + std::to_string(herr[1]) + ")");
        return AMR_ERR_MATERIAL;
    }
    if (herr[3]) {
        amr_set_error("solve_system!: NaN values in internal forces vector");
        return AMR_ERR_NAN;
    }
    return AMR_OK;
}

// Predictor and corrector of the :ME, :BE and :Ralston schemes
// (src/mech/mech-solver.jl:280-288, 325-349): ΔUitr = p1 ΔUi, Rtr = q11 R;
// solve_system!(K, ΔUitr, Rtr); update_state!(ΔUa + ΔUitr) from the
// committed state; K2 = mount_K at that state; K = a1 K + a2 K2;
// solve_system!(K, ΔUi, R); then amr_nr_finish's update_state!(ΔUa + ΔUi),
// ΔUa += ΔUi and residual.  The predictor's update must succeed (its
// failure ends the iteration as the reference's syserror does).
static int nr_iteration_pc(amr_model *m, int it, const amr_solve_opts *o, amr_nr_info *info)
{
    static const double P[4][4] = {{1.0, 1.0, 0.0, 0.0}, {1.0, 1.0, 0.5, 0.5}, {1.0, 1.0, 0.0, 1.0},
                                   {2.0 / 3.0, 2.0 / 3.0, 0.25, 0.75}};
    const double p1 = P[m->scheme][0], q11 = P[m->scheme][1], a1 = P[m->scheme][2], a2 = P[m->scheme][3];
    double **v = m->sv;
    double *dUitr = m->v[11], *Rtr = m->v[7];
    amr_nr_info inf{};
    auto fail = [&](int rc) {
        inf.status = rc;
        if (info) *info = inf;
        return rc;
    };
    // predictor: mount_K at the current state, the scaled system
    if (it > 1) TRY(amr_vec_fill(m, v[SV_DUI], 0.0));
    AMR_HIP(hipMemsetAsync(m->err, 0x7f, 3 * sizeof(int32_t), m->stream));
    TRY(amr_launch_ke(m));
    TRY(amr_launch_assemble(m, false));
    m->assembled = true;
    TRY(amr_vec_scale(m, p1, v[SV_DUI], dUitr));
    TRY(amr_vec_scale(m, q11, v[SV_R], Rtr));
    amr_solve_info si{};
    if (m->nu > 0) {
        TRY(amr_vec_solve_prep(m, dUitr, Rtr, m->v[8], m->v[9], m->v[10]));
        const int rc = amr_pcg_solve(m, m->v[10], v[SV_X], o, &si);
        inf.pcg_iterations = si.iterations;
        inf.replacements = si.refinements;
        if (rc != AMR_OK) return rc == AMR_ERR_SOLVE ? fail(rc) : rc;
        TRY(amr_vec_combine(m, v[SV_X], dUitr));
    }
    TRY(amr_vec_add(m, v[SV_DUA], dUitr, v[SV_DUT]));
    TRY(amr_launch_update(m, v[SV_DUT], v[SV_DFIN]));
    TRY(amr_vec_nan_flag(m, v[SV_DFIN], m->err + 3));
    if (int rc = update_status(m)) return fail(rc);
    // corrector: the predictor's tangent kept, mount_K at the predicted state
    if (!m->Kpred) TRY(dalloc(m, &m->Kpred, (size_t)kv_size(m->pat.nblk)));
    AMR_HIP(hipMemcpyAsync(m->Kpred, m->Kv, sizeof(double) * kv_size(m->pat.nblk), hipMemcpyDeviceToDevice,
                           m->stream));
    AMR_HIP(hipMemsetAsync(m->err, 0x7f, 3 * sizeof(int32_t), m->stream));
    TRY(amr_launch_ke(m));
    TRY(amr_launch_assemble(m, false));
    TRY(amr_launch_combine_kk(m, a1, m->Kpred, a2));
    amr_solve_info si2{};
    if (m->nu > 0) {
        TRY(amr_vec_solve_prep(m, v[SV_DUI], v[SV_R], m->v[8], m->v[9], m->v[10]));
        const int rc = amr_pcg_solve(m, m->v[10], v[SV_X], o, &si2);
        if (rc != AMR_OK && rc != AMR_ERR_SOLVE) return rc;
    }
    const int rc = amr_nr_finish(m, m->nu > 0 ? &si2 : nullptr, info);
    if (info) {
        info->pcg_iterations += si.iterations;
        info->replacements += si.refinements;
    }
    return rc;
}

extern "C" int amr_nr_iteration(amr_model *m, int it, const amr_solve_opts *opts, amr_nr_info *info)
{
    PhaseRange range_("NR iteration");
    amr_solve_opts o = opts ? *opts : default_opts();
    if (m && m->scheme != AMR_SCHEME_FE) {
        if (!m->sv[SV_U] || it < 1) return AMR_ERR_ARG;
        if (amr_pool_is_member(m)) {
            amr_set_error("the model is solving in a pool; its matrix and solver vectors are in use");
            return AMR_ERR_ARG;
        }
        AMR_HIP(hipSetDevice(m->device));
        return nr_iteration_pc(m, it, &o, info);
    }
    TRY(amr_nr_prepare(m, it));
    amr_solve_info si{};
    if (m->nu > 0) {
        const int rc = amr_pcg_solve(m, m->v[10], m->sv[SV_X], &o, &si);
        if (rc != AMR_OK && rc != AMR_ERR_SOLVE) return rc;
    }
    return amr_nr_finish(m, m->nu > 0 ? &si : nullptr, info);
}

extern "C" int amr_increment_commit(amr_model *m)
{
    if (!m || !m->sv[SV_U]) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    TRY(amr_vec_commit(m));
    return amr_state_commit(m);
}

extern "C" int amr_increment_reject(amr_model *m)
{
    if (!m) return AMR_ERR_ARG;
    return amr_state_restore(m);
}

extern "C" int amr_stage_get(amr_model *m, double *U, double *F, int last)
{
    if (!m || !m->sv[SV_U]) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    const size_t n = (size_t)m->nnodes * 3;
    std::lock_guard<std::mutex> lp(g_pin_mu);
    double *pin;
    TRY(pinned_buffer(m->device, 2 * n, &pin));
    TRY(d2h(pin, m->sv[last ? SV_DUA : SV_U], n, m->stream));
    TRY(d2h(pin + n, m->sv[last ? SV_DFIN : SV_F], n, m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    if (U) node_to_eq(m, pin, U);
    if (F) node_to_eq(m, pin + n, F);
    return AMR_OK;
}

extern "C" int amr_timing(amr_model *m, int enable, double *spmv_ms, int64_t *spmv_launches)
{
    if (!m) return AMR_ERR_ARG;
    if (enable && !m->timing) {
        m->t_spmv_ms = 0.0;
        m->n_spmv = 0;
    }
    m->timing = enable != 0;
    if (spmv_ms) *spmv_ms = m->t_spmv_ms;
    if (spmv_launches) *spmv_launches = m->n_spmv;
    return AMR_OK;
}

extern "C" int amr_set_spmv_lanes(amr_model *m, int lanes_fp64, int lanes_fp32)
{
    auto ok = [](int l) { return l == 0 || l == 8 || l == 16 || l == 32 || l == 64; };
    if (!m || !ok(lanes_fp64) || !ok(lanes_fp32)) {
        amr_set_error("amr_set_spmv_lanes: lanes must be 0, 8, 16, 32 or 64");
        return AMR_ERR_ARG;
    }
    m->lanes64 = lanes_fp64;
    m->lanes32 = m->lanes32d = lanes_fp32;
    if (!lanes_fp64 || !lanes_fp32) m->lanes64 = m->lanes32 = m->lanes32d = 0;
    return AMR_OK;
}

extern "C" int amr_tune_spmv_lanes(amr_model *m)
{
    if (!m) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    return amr_tune_lanes(m);
}

extern "C" int amr_get_spmv_lanes(amr_model *m, int *out)
{
    if (!m || !out) return AMR_ERR_ARG;
    out[0] = m->lanes64;
    out[1] = m->lanes32;
    out[2] = m->lanes32d;
    return AMR_OK;
}

// ---------------------------------------------------------------- pools

extern "C" int amr_pool_create(amr_pool **out, int device, int max_members, int check_every)
{
    if (!out) return AMR_ERR_ARG;
    return amr_pool_new(out, device, max_members, check_every, nullptr);
}

extern "C" int amr_pool_destroy(amr_pool *p)
{
    amr_pool_free(p);
    return AMR_OK;
}

extern "C" int amr_pool_attach(amr_pool *p, amr_model *m)
{
    if (!p || !m) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    return amr_pool_attach_model(p, m);
}

extern "C" int amr_pool_detach(amr_model *m)
{
    if (!m) return AMR_ERR_ARG;
    AMR_HIP(hipSetDevice(m->device));
    return amr_pool_detach_model(m);
}

extern "C" int amr_pool_add(amr_pool *p, amr_model *m, const amr_solve_opts *opts)
{
    if (!p || !m || !m->sv[SV_U]) return AMR_ERR_ARG;
    if (m->nu == 0) {
        amr_set_error("amr_pool_add: the model has no unknowns; call amr_nr_finish with si = NULL");
        return AMR_ERR_ARG;
    }
    if (m->scheme != AMR_SCHEME_FE) {
        amr_set_error("amr_pool_add: pooled solves run scheme :FE; use amr_nr_iteration for the others");
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipSetDevice(m->device));
    const amr_solve_opts o = opts ? *opts : default_opts();
    return amr_pool_join(p, m, m->v[10], m->sv[SV_X], &o);
}

static int copy_results(const std::vector<amr_pool_result> &res, int *nfinished, amr_model **models,
                        amr_solve_info *infos)
{
    *nfinished = (int)res.size();
    for (size_t i = 0; i < res.size(); i++) {
        models[i] = res[i].m;
        infos[i] = res[i].info;
    }
    return AMR_OK;
}

extern "C" int amr_pool_step(amr_pool *p, int capacity, int *nfinished, amr_model **models, amr_solve_info *infos)
{
    PhaseRange range_("pool step");
    if (!p || !nfinished) return AMR_ERR_ARG;
    if (capacity < amr_pool_count(p) || (capacity > 0 && (!models || !infos))) {
        amr_set_error("amr_pool_step: the output arrays must hold every member of the pool");
        return AMR_ERR_ARG;
    }
    std::vector<amr_pool_result> res;
    const int rc = amr_pool_run(p, res);
    copy_results(res, nfinished, models, infos);
    return rc;
}

extern "C" int amr_pool_step_begin(amr_pool *p)
{
    PhaseRange range_("pool step (queue)");
    if (!p) return AMR_ERR_ARG;
    return amr_pool_step_launch(p);
}

extern "C" int amr_pool_step_done(amr_pool *p, int *done)
{
    if (!p || !done) return AMR_ERR_ARG;
    return amr_pool_step_query(p, done);
}

extern "C" int amr_pool_release_left(amr_pool *p)
{
    if (!p) return AMR_ERR_ARG;
    return amr_pool_hand_back(p);
}

extern "C" int amr_pool_step_end(amr_pool *p, int capacity, int *nfinished, amr_model **models,
                                 amr_solve_info *infos)
{
    if (!p || !nfinished) return AMR_ERR_ARG;
    if (capacity < amr_pool_step_members(p) || (capacity > 0 && (!models || !infos))) {
        amr_set_error("amr_pool_step_end: the output arrays must hold every member of the step");
        return AMR_ERR_ARG;
    }
    std::vector<amr_pool_result> res;
    const int rc = amr_pool_step_collect(p, res);
    copy_results(res, nfinished, models, infos);
    return rc;
}

extern "C" int amr_pool_size(amr_pool *p, int *members)
{
    if (!p || !members) return AMR_ERR_ARG;
    *members = amr_pool_count(p);
    return AMR_OK;
}

extern "C" int amr_pool_timing(amr_pool *p, int enable, double *spmv_ms, int64_t *spmv_launches)
{
    if (!p) return AMR_ERR_ARG;
    return amr_pool_set_timing(p, enable, spmv_ms, spmv_launches);
}

extern "C" int amr_set_ke_path(amr_model *m, int path)
{
    if (!m || (path != AMR_KE_EXACT && path != AMR_KE_MFMA)) {
        amr_set_error("amr_set_ke_path: path must be AMR_KE_EXACT or AMR_KE_MFMA");
        return AMR_ERR_ARG;
    }
    if (path == AMR_KE_MFMA && m->etype != AMR_HEX20 && !m->all_elastic) {
        amr_set_error("amr_set_ke_path: the matrix-core path takes HEX20 elements or LinearElastic models");
        return AMR_ERR_ARG;
    }
    m->ke_path = path;
    return AMR_OK;
}

extern "C" int amr_set_elem_tiles(amr_model *m, int ach, int epw)
{
    if (!m) return AMR_ERR_ARG;
    return amr_elem_tiles_set(m, ach, epw);
}

extern "C" int amr_get_elem_tiles(amr_model *m, int *ach, int *epw)
{
    if (!m || !ach || !epw) return AMR_ERR_ARG;
    return amr_elem_tiles(m, ach, epw);
}

extern "C" int amr_tune_elem_tiles(amr_model *m, int *ach, int *epw)
{
    if (!m || !ach || !epw) return AMR_ERR_ARG;
    if (amr_pool_is_member(m)) {
        amr_set_error("the model is solving in a pool; its matrix and solver vectors are in use");
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipSetDevice(m->device));
    return amr_elem_tiles_tune(m, ach, epw);
}

extern "C" int amr_pool_set_graphs(amr_pool *p, int enable, int64_t *captures, int64_t *replays)
{
    if (!p) return AMR_ERR_ARG;
    return amr_pool_graphs(p, enable, captures, replays);
}

extern "C" int amr_pool_stats(amr_pool *p, int64_t *out)
{
    if (!p || !out) return AMR_ERR_ARG;
    return amr_pool_counts(p, out);
}

extern "C" int amr_model_device_bytes(amr_model *m, int64_t *bytes)
{
    if (!m || !bytes) return AMR_ERR_ARG;
    *bytes = m->dev_bytes;
    return AMR_OK;
}

extern "C" int amr_mem_info(int device, int64_t *free_bytes, int64_t *total_bytes)
{
    AMR_HIP(hipSetDevice(device));
    size_t f = 0, t = 0;
    AMR_HIP(hipMemGetInfo(&f, &t));
    {
        // blocks the library holds for reuse are free for models
        std::lock_guard<std::mutex> lk(g_dmem_mu);
        auto it = g_arena.find(device);
        if (it != g_arena.end()) f += it->second.bytes;
    }
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
    return AMR_OK;
}

extern "C" int amr_set_coarse(amr_model *m, int target_nodes, int max_aggregates)
{
    if (!m || target_nodes < 1 || max_aggregates < 1 || max_aggregates > 340) {
        amr_set_error("amr_set_coarse: target_nodes >= 1 and 1 <= max_aggregates <= 340");
        return AMR_ERR_ARG;
    }
    if (m->coarse_built) {
        amr_set_error("amr_set_coarse: the coarse space is already built");
        return AMR_ERR_ARG;
    }
    m->coarse_target = target_nodes;
    m->coarse_max = max_aggregates;
    AMR_HIP(hipSetDevice(m->device));
    return amr_coarse_build(m);
}

extern "C" int amr_set_coarse_refresh(amr_model *m, int policy)
{
    if (!m || policy < AMR_COARSE_PER_DOF_MAP || policy > AMR_COARSE_PER_ASSEMBLY) {
        amr_set_error("amr_set_coarse_refresh: policy must be AMR_COARSE_PER_DOF_MAP, _PER_INCREMENT or _PER_ASSEMBLY");
        return AMR_ERR_ARG;
    }
    m->coarse_refresh = policy;
    return AMR_OK;
}

extern "C" int amr_get_coarse(amr_model *m, int64_t *out)
{
    if (!m || !out) return AMR_ERR_ARG;
    out[0] = m->coarse_built ? m->nagg : 0;
    out[1] = m->coarse_built ? m->ncoarse : 0;
    return AMR_OK;
}

extern "C" int amr_get_coarse_space(amr_model *m, int32_t *agg, double *kc_inv)
{
    if (!m) return AMR_ERR_ARG;
    if (!m->coarse_built) {
        amr_set_error("amr_get_coarse_space: no two-level solve has built the coarse space yet");
        return AMR_ERR_ARG;
    }
    AMR_HIP(hipSetDevice(m->device));
    if (agg) TRY(d2h(agg, m->c_agg, (size_t)m->nnodes, m->stream));
    if (kc_inv) TRY(d2h(kc_inv, m->Kinv, (size_t)m->ncoarse * m->ncoarse, m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    return AMR_OK;
}
