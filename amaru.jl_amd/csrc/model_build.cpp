// Host-side construction of the node-block sparsity structure and of the
// deterministic gather maps used by assembly and by the internal-force sum.
//
// K is stored as 3x3 blocks, one block row per node, block columns sorted by
// node id.  A block (A,B) exists when nodes A and B share an element; the
// scalar pattern the reference builds (src/mech/mech-solver.jl:86-102, drop
// |Ke_ij| < eps) is a subset of it and is recovered from per-scalar survivor
// flags written at assembly.
//
// Gather maps replace the reference's triplet push!/sparse() and its
// ΔFin[map] .+= ΔF scatter (src/mech/mech-solver.jl:129-136): for every
// block, the list of element blocks (e, la, lb) that land in it, in
// ascending element order; for every node, the list of (e, la).  Summing
// each list in order reproduces the reference's element-order sums without
// atomics.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "amr_internal.h"

int amr_build_pattern(int nn, int64_t nnodes, int64_t ne, const int32_t *conn, AmrPattern &P)
{
    P.nnodes = nnodes;
    // node -> elements incidence, ascending element order
    std::vector<int32_t> ninc(nnodes + 1, 0);
    for (int64_t e = 0; e < ne; e++)
        for (int a = 0; a < nn; a++) {
            int32_t n = conn[e * nn + a];
            if (n < 0 || n >= nnodes) {
                amr_set_error("connectivity references node " + std::to_string(n) + " outside [0, " +
                              std::to_string(nnodes) + ")");
                return AMR_ERR_ARG;
            }
            ninc[n + 1]++;
        }
    for (int64_t n = 0; n < nnodes; n++) ninc[n + 1] += ninc[n];
    std::vector<int32_t> inc_e(ninc[nnodes]), inc_a(ninc[nnodes]);
    {
        std::vector<int32_t> pos(ninc.begin(), ninc.end() - 1);
        for (int64_t e = 0; e < ne; e++)
            for (int a = 0; a < nn; a++) {
                int32_t n = conn[e * nn + a];
                inc_e[pos[n]] = (int32_t)e;
                inc_a[pos[n]] = a;
                pos[n]++;
            }
    }
    // force gather lists: exactly the incidence lists (ascending e)
    P.fptr.assign(ninc.begin(), ninc.end());
    P.fsrc.resize(ninc[nnodes]);
    for (size_t i = 0; i < P.fsrc.size(); i++) P.fsrc[i] = inc_e[i] * nn + inc_a[i];

    // block rows: sorted unique neighbours of each node (a stamp per node
    // drops repeats before the sort)
    P.rowptr.assign(nnodes + 1, 0);
    P.col.clear();
    P.col.reserve((size_t)nnodes * 32);
    P.diag.assign(nnodes, -1);
    std::vector<int32_t> stamp(nnodes, -1), buf;
    for (int64_t n = 0; n < nnodes; n++) {
        buf.clear();
        for (int32_t i = ninc[n]; i < ninc[n + 1]; i++) {
            const int32_t *el = conn + (int64_t)inc_e[i] * nn;
            for (int b = 0; b < nn; b++)
                if (stamp[el[b]] != (int32_t)n) {
                    stamp[el[b]] = (int32_t)n;
                    buf.push_back(el[b]);
                }
        }
        std::sort(buf.begin(), buf.end());
        auto it = std::lower_bound(buf.begin(), buf.end(), (int32_t)n);
        if (it != buf.end() && *it == n) P.diag[n] = (int32_t)(P.col.size() + (it - buf.begin()));
        P.col.insert(P.col.end(), buf.begin(), buf.end());
        P.rowptr[n + 1] = (int32_t)P.col.size();
    }
    P.nblk = P.rowptr[nnodes];
    if ((int64_t)ne * nn * nn > INT32_MAX) {
        amr_set_error("model too large for 32-bit element-block ids");
        return AMR_ERR_ARG;
    }

    // block contribution lists, one block row at a time: the element blocks
    // (e, a, b) of row A come from A's incident elements, visited in
    // ascending e, and each element adds at most one entry to a block, so
    // every list comes out sorted by element.  Pass 1 counts, pass 2 fills.
    std::vector<int32_t> &where = stamp;  // column -> block index within the current row
    P.cptr.assign(P.nblk + 1, 0);
    for (int64_t A = 0; A < nnodes; A++) {
        for (int32_t k = P.rowptr[A]; k < P.rowptr[A + 1]; k++) where[P.col[k]] = k;
        for (int32_t i = ninc[A]; i < ninc[A + 1]; i++) {
            const int32_t *el = conn + (int64_t)inc_e[i] * nn;
            for (int b = 0; b < nn; b++) P.cptr[where[el[b]] + 1]++;
        }
    }
    for (int64_t k = 0; k < P.nblk; k++) P.cptr[k + 1] += P.cptr[k];
    P.ncontrib = P.cptr[P.nblk];
    P.csrc.resize(P.ncontrib);
    std::vector<int64_t> fill;
    for (int64_t A = 0; A < nnodes; A++) {
        const int32_t r0 = P.rowptr[A];
        fill.assign(P.cptr.begin() + r0, P.cptr.begin() + P.rowptr[A + 1]);
        for (int32_t k = r0; k < P.rowptr[A + 1]; k++) where[P.col[k]] = k - r0;
        for (int32_t i = ninc[A]; i < ninc[A + 1]; i++) {
            const int64_t e = inc_e[i];
            const int32_t *el = conn + e * nn;
            const int32_t base = (int32_t)((e * nn + inc_a[i]) * nn);
            for (int b = 0; b < nn; b++) P.csrc[fill[where[el[b]]]++] = base + b;
        }
    }
    return AMR_OK;
}

// ---------------------------------------------------------------------------
// Coarse space of the two-level preconditioner: aggregates of nodes by
// recursive coordinate bisection (split the longest extent at the median of
// (coordinate, node id), until every part has at most `target` nodes; a
// linear-time selection per split), and the index lists that let the device
// form Kc = Phi^T K Phi deterministically:
//   pairs:  for each block row i, the distinct aggregates b of its columns
//           (ascending), one output slot per (i, b);
//   cblocks: for each coarse block (a, b), the slots of rows i in a with
//           aggregate b, ascending i.
int amr_build_coarse(int64_t nnodes, const double *X, const int32_t *rowptr, const int32_t *col, int target,
                     AmrCoarse &C)
{
    if (target < 1) target = 1;
    C = AmrCoarse();
    C.agg.assign(nnodes, 0);
    // the nodes with their coordinates, permuted in place by the splits
    struct Pt {
        double x[3];
        int32_t id;
    };
    std::vector<Pt> pts(nnodes);
    for (int64_t i = 0; i < nnodes; i++) pts[i] = Pt{{X[i * 3], X[i * 3 + 1], X[i * 3 + 2]}, (int32_t)i};
    // explicit stack of [lo, hi) ranges of pts; parts are numbered in the
    // order they are finished (depth first, lower half first)
    std::vector<std::pair<int64_t, int64_t>> st{{0, nnodes}};
    int32_t nagg = 0;
    while (!st.empty()) {
        auto [lo, hi] = st.back();
        st.pop_back();
        if (hi - lo <= target) {
            for (int64_t k = lo; k < hi; k++) C.agg[pts[k].id] = nagg;
            nagg++;
            continue;
        }
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t k = lo; k < hi; k++)
            for (int d = 0; d < 3; d++) {
                mn[d] = std::min(mn[d], pts[k].x[d]);
                mx[d] = std::max(mx[d], pts[k].x[d]);
            }
        int ax = 0;
        for (int d = 1; d < 3; d++)
            if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
        const int64_t mid = lo + (hi - lo) / 2;
        // the lower half is the (mid - lo) smallest by (coordinate, id), a
        // set that does not depend on the order inside [lo, hi)
        std::nth_element(pts.begin() + lo, pts.begin() + mid, pts.begin() + hi, [ax](const Pt &a, const Pt &b) {
            return a.x[ax] < b.x[ax] || (a.x[ax] == b.x[ax] && a.id < b.id);
        });
        st.push_back({mid, hi});
        st.push_back({lo, mid});
    }
    C.nagg = nagg;
    // nodes of each aggregate, ascending; centroids
    C.aggptr.assign(nagg + 1, 0);
    for (int64_t i = 0; i < nnodes; i++) C.aggptr[C.agg[i] + 1]++;
    for (int a = 0; a < nagg; a++) C.aggptr[a + 1] += C.aggptr[a];
    C.aggnodes.assign(nnodes, 0);
    {
        std::vector<int32_t> fill(C.aggptr.begin(), C.aggptr.end() - 1);
        for (int64_t i = 0; i < nnodes; i++) C.aggnodes[fill[C.agg[i]]++] = (int32_t)i;
    }
    C.cent.assign((size_t)nagg * 3, 0.0);
    for (int a = 0; a < nagg; a++) {
        double s[3] = {0.0, 0.0, 0.0};
        for (int32_t k = C.aggptr[a]; k < C.aggptr[a + 1]; k++)
            for (int d = 0; d < 3; d++) s[d] += X[(int64_t)C.aggnodes[k] * 3 + d];
        const double w = 1.0 / (double)(C.aggptr[a + 1] - C.aggptr[a]);
        for (int d = 0; d < 3; d++) C.cent[a * 3 + d] = s[d] * w;
    }
    // (row, aggregate) pairs: the distinct aggregates of each row's columns
    // (a stamp per aggregate drops repeats), ascending
    C.pptr.assign(nnodes + 1, 0);
    C.prow.clear();
    std::vector<int32_t> seen, stamp(nagg, -1);
    for (int64_t i = 0; i < nnodes; i++) {
        seen.clear();
        for (int32_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
            const int32_t b = C.agg[col[k]];
            if (stamp[b] != (int32_t)i) {
                stamp[b] = (int32_t)i;
                seen.push_back(b);
            }
        }
        std::sort(seen.begin(), seen.end());
        for (int32_t b : seen) {
            C.pagg.push_back(b);
            C.prow.push_back((int32_t)i);
        }
        C.pptr[i + 1] = (int32_t)C.pagg.size();
    }
    // coarse blocks (a, b) ascending, slots ascending in row: a stable
    // counting sort by b, then by a (slots are generated in row order)
    const int64_t np = (int64_t)C.pagg.size();
    std::vector<int64_t> key(np);
    for (int64_t i = 0; i < nnodes; i++)
        for (int32_t s = C.pptr[i]; s < C.pptr[i + 1]; s++) key[s] = (int64_t)C.agg[i] * nagg + C.pagg[s];
    std::vector<int32_t> ord(np), tmp(np);
    {
        std::vector<int64_t> cnt(nagg + 1);
        auto pass = [&](const std::vector<int32_t> &in, std::vector<int32_t> &out, bool by_a) {
            std::fill(cnt.begin(), cnt.end(), 0);
            for (int64_t t = 0; t < np; t++) {
                const int64_t k = key[in[t]];
                cnt[(by_a ? k / nagg : k % nagg) + 1]++;
            }
            for (int32_t v = 0; v < nagg; v++) cnt[v + 1] += cnt[v];
            for (int64_t t = 0; t < np; t++) {
                const int64_t k = key[in[t]];
                out[cnt[by_a ? k / nagg : k % nagg]++] = in[t];
            }
        };
        for (int64_t t = 0; t < np; t++) tmp[t] = (int32_t)t;
        pass(tmp, ord, false);
        pass(ord, tmp, true);
        ord.swap(tmp);
    }
    C.cbptr.push_back(0);
    for (int64_t t = 0; t < np; t++) {
        if (t > 0 && key[ord[t]] != key[ord[t - 1]]) C.cbptr.push_back((int32_t)t);
        if (t == 0 || key[ord[t]] != key[ord[t - 1]]) {
            C.cba.push_back((int32_t)(key[ord[t]] / nagg));
            C.cbb.push_back((int32_t)(key[ord[t]] % nagg));
        }
        C.cbslot.push_back(ord[t]);
    }
    C.cbptr.push_back((int32_t)np);
    return AMR_OK;
}
