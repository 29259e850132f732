/*
 * firmcore_oracle.c -- CPU restatement of the reference's FirmCore algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this code, and only as the checker.
 * The product (pargridcpu_amd/, libhipgrid.so) never links or calls it.
 *
 * Every function restates one reference routine in plain C and cites the
 * file:line it follows.  Graph input is per-layer symmetric CSR:
 * row_ptr[l][0..n] (int64) and col[l][...] (uint32), rows sorted ascending,
 * no duplicates, no self-loops.  The reference's degree-prefixed row
 * adj_lst[v] = [deg, nbrs...] (Graph/Graph.cpp:17-50) maps to
 * deg = row_ptr[l][v+1] - row_ptr[l][v], nbrs = col[l] + row_ptr[l][v].
 *
 * Core index output is core_index[lmd-1][v] (int32, lambda-major):
 *   index_lmd(v) = max{k >= 1 : v in S(k, lmd)}, or 0 if v is not in S(1, lmd),
 * where S(k, lmd) is the (k, lmd)-FirmCore.  Cell lengths are
 * cell_len[(lmd-1)*cap + k] = |S(k, lmd)| for k = 0..cap-1 (k = 0 is n).
 *
 * How each function is pinned (checked by tests/test_oracle.py):
 *   orc_pargrid, orc_tree (lpe and kpe), orc_coreindex: against the golden
 *     vectors written by the reference built from source (tests/golden/),
 *     against each other, against orc_bruteforce on small graphs, and against
 *     closed forms (L = 1 k-core numbers, cliques, stars, cycles, paths,
 *     identical layers).
 *   orc_coreindex_lmd, orc_coreindex_merged_lmd, orc_merge: orc_coreindex is
 *     orc_merge followed by the loop of orc_coreindex_merged_lmd over lmd,
 *     so every pin of orc_coreindex covers them; the golden tests also run
 *     them from threads, one lmd each.
 *   orc_check_lmd: accepts every golden core index written by the reference,
 *     and rejects each planted off-by-one (+1 or -1 at one vertex) of them.
 *   orc_bruteforce: definition-level fixed point, checked by the closed forms.
 *   orc_wds: against the reference's printed WdsList result (golden), plus a
 *     brute-force per-layer induced edge count on small graphs.
 *   orc_wds_coreindex: against get_wds's per-lmd results from the reference
 *     built from source (golden), plus an independent numpy recomputation.
 */
#define _DEFAULT_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>

typedef struct {
    int64_t n;
    int32_t L;
    const int64_t* const* row_ptr;
    const uint32_t* const* col;
} orc_graph;

static inline int64_t deg_of(const orc_graph* g, int l, int64_t v) {
    return g->row_ptr[l][v + 1] - g->row_ptr[l][v];
}

/* ------------------------------------------------------------------ ParGrid
 * Restates FCSyncLeft::Execute / BuildSubFCTreeSync / PathSerialSync /
 * PathByK / PeelSync (CoreParallel/FCSyncLeft.cpp:40-252) single-threaded:
 * the same two phases per cell (scan every vertex, then drain the removal
 * queue with per-layer decrements), the same lambda-spine at k = 1 followed
 * by one k-chain per spine node started from that node's saved state.
 * A vertex removed while peeling cell (k, lmd) gets index k - 1. */

/* PeelSync (FCSyncLeft.cpp:40-131).  Returns the number removed. */
static int64_t pargrid_peel(const orc_graph* g, int32_t* degs, uint8_t* valid, int32_t* cnts,
                            int32_t* buff, uint32_t k, uint32_t lmd, int32_t* idx_row) {
    const int64_t n = g->n;
    const int L = g->L;
    int64_t end = 0, start = 0;
    /* phase 1 (FCSyncLeft.cpp:65-83) */
    for (int64_t v = 0; v < n; ++v) {
        if (!valid[v]) { cnts[v] = 0; continue; }
        int32_t cnt = 0;
        for (int l = 0; l < L; ++l) cnt += degs[v * L + l] >= (int32_t)k;
        if (cnt < (int32_t)lmd) {
            cnts[v] = 0;
            valid[v] = 0;
            buff[end++] = (int32_t)v;
        } else {
            cnts[v] = cnt;
        }
    }
    /* phase 2 (FCSyncLeft.cpp:91-108) */
    while (start < end) {
        const int32_t vv = buff[start++];
        if (idx_row) idx_row[vv] = (int32_t)k - 1;
        for (int l = 0; l < L; ++l) {
            const uint32_t* nb = g->col[l] + g->row_ptr[l][vv];
            const int64_t d = deg_of(g, l, vv);
            for (int64_t i = 0; i < d; ++i) {
                const uint32_t u = nb[i];
                const int32_t old = degs[(int64_t)u * L + l]--;
                if (old == (int32_t)k) {
                    const int32_t oc = cnts[u]--;
                    if (oc == (int32_t)lmd && valid[u]) {
                        valid[u] = 0;
                        buff[end++] = (int32_t)u;
                    }
                }
            }
        }
    }
    return end;
}

int orc_pargrid(int64_t n, int32_t L, const int64_t* const* row_ptr, const uint32_t* const* col,
                int32_t* core_index, int64_t* cell_len, int32_t cap, int32_t* kmax) {
    orc_graph g = {n, L, row_ptr, col};
    int32_t* degs = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n * L + 1));
    int32_t* degs2 = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n * L + 1));
    uint8_t* valid = (uint8_t*)malloc((size_t)n + 1);
    uint8_t* valid2 = (uint8_t*)malloc((size_t)n + 1);
    int32_t* cnts = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int32_t* buff = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    if (!degs || !degs2 || !valid || !valid2 || !cnts || !buff) {
        free(degs); free(degs2); free(valid); free(valid2); free(cnts); free(buff);
        return -3;
    }
    /* FCSyncLeft.cpp:218-235 */
    for (int64_t v = 0; v < n; ++v)
        for (int l = 0; l < L; ++l) degs[v * L + l] = (int32_t)deg_of(&g, l, v);
    memset(valid, 1, (size_t)n);
    memset(core_index, 0, sizeof(int32_t) * (size_t)(n * L));
    if (cell_len) {
        memset(cell_len, 0, sizeof(int64_t) * (size_t)L * (size_t)cap);
        if (cap > 0) for (int l = 0; l < L; ++l) cell_len[(size_t)l * cap] = n;
    }
    for (int l = 0; l < L; ++l) kmax[l] = 0;
    int64_t total = 0;
    /* spine (PathSerialSync, FCSyncLeft.cpp:134-152) interleaved with the
     * k-chain of each spine node (PathByK, :155-170, run from the node's
     * saved degs/valid/total as BuildSubFCTreeSync does at :191-198). */
    for (uint32_t lmd = 1; lmd <= (uint32_t)L; ++lmd) {
        total += pargrid_peel(&g, degs, valid, cnts, buff, 1, lmd, NULL);
        int64_t length = n - total;
        if (length <= 0) break;
        int32_t* idx_row = core_index + (size_t)(lmd - 1) * (size_t)n;
        if (cell_len && cap > 1) cell_len[(size_t)(lmd - 1) * cap + 1] = length;
        memcpy(degs2, degs, sizeof(int32_t) * (size_t)(n * L));
        memcpy(valid2, valid, (size_t)n);
        int64_t t2 = total;
        uint32_t k = 2;
        for (;;) {
            t2 += pargrid_peel(&g, degs2, valid2, cnts, buff, k, lmd, idx_row);
            const int64_t len2 = n - t2;
            if (len2 <= 0) break;
            if (cell_len && (int32_t)k < cap) cell_len[(size_t)(lmd - 1) * cap + k] = len2;
            ++k;
        }
        kmax[lmd - 1] = (int32_t)k - 1;
    }
    free(degs); free(degs2); free(valid); free(valid2); free(cnts); free(buff);
    return 0;
}

/* --------------------------------------------------------------- lpe / kpe
 * Restates FCTreeBuilderLeft::{Execute, BuildSubFCTree, peel, check,
 * restore, PrintCoreInfor} (Core/FCTreeBuilderLeft.cpp:8-285) and, with
 * right_spine = 1, FCTreeBuilderRight (Core/FCTreeBuilderRight.cpp:8-245):
 * one recursive DFS over the fcgrid with the in-place core[]/pos[] partition
 * (removed vertices in core[0..e), alive in core[e..n)) and degree restore
 * after each subtree.  lpe: left child (k+1, lmd), right child (1, lmd+1)
 * only at k = 1.  kpe: right child (k, lmd+1), left child (k+1, 1) only at
 * lmd = 1.
 *
 * With wds != 0 every visited cell is scored as FCPathLevelLeft::
 * constructCore does (Core/FCPathLevelLeft.cpp:57-99) and the running max
 * is kept as FCTree::WeightDenestSubgraph + main.cpp:402-424 keep it (a
 * float max updated with >= in pre-order, which for lpe is the visiting
 * order: lambda-major, k ascending).  Only meaningful for lpe.
 */
typedef struct {
    const orc_graph* g;
    int32_t* degs;
    uint32_t* core;
    uint32_t* pos;
    int32_t* core_index;
    int64_t* cell_len;
    int32_t cap;
    int32_t* kmax;
    int right_spine;
    int64_t count;
    int64_t sum_num_of_edge_fc;
    int64_t* edge_fc_k1; /* lpe's printed num_of_edge_fc for (k, 1), index k */
    int32_t edge_fc_cap;
    /* WDS */
    int wds;
    const double* pow_tab; /* pow(l+1, beta), l = 0..L-1 */
    float max_density;
    uint32_t res_k, res_lmd, res_len;
    double* cell_density; /* optional [L][cap] per-cell layer_density */
    int32_t* cell_sel;    /* optional [L][cap] num_select_layer */
} tree_ctx;

static int tree_check(const tree_ctx* c, uint32_t u, uint32_t k, uint32_t lmd) {
    const int L = c->g->L;
    uint32_t cnt = 0;
    for (int l = 0; l < L; ++l) {
        if (c->degs[(int64_t)u * L + l] >= (int32_t)k) {
            if (++cnt >= lmd) return 1;
        }
    }
    return 0;
}

static uint32_t tree_peel(tree_ctx* c, uint32_t k, uint32_t lmd, uint32_t s, uint32_t e) {
    const orc_graph* g = c->g;
    const int L = g->L;
    while (s < e) {
        const uint32_t old_e = e;
        for (int i = 0; i < L; ++i) {
            for (uint32_t j = s; j < old_e; ++j) {
                const uint32_t v = c->core[j];
                const uint32_t* nb = g->col[i] + g->row_ptr[i][v];
                const int64_t d = deg_of(g, i, v);
                for (int64_t t = 0; t < d; ++t) {
                    const uint32_t u = nb[t];
                    c->degs[(int64_t)u * L + i]--;
                    if (c->degs[(int64_t)u * L + i] < (int32_t)k && c->pos[u] >= e) {
                        if (!tree_check(c, u, k, lmd)) {
                            const uint32_t pu = c->pos[u];
                            c->core[pu] = c->core[e];
                            c->core[e] = u;
                            c->pos[c->core[pu]] = pu;
                            c->pos[u] = e;
                            ++e;
                        }
                    }
                }
            }
        }
        s = old_e;
    }
    return e;
}

static void tree_restore(tree_ctx* c, uint32_t old_e, uint32_t new_e) {
    const orc_graph* g = c->g;
    const int L = g->L;
    for (int i = 0; i < L; ++i)
        for (uint32_t j = old_e; j < new_e; ++j) {
            const uint32_t v = c->core[j];
            const uint32_t* nb = g->col[i] + g->row_ptr[i][v];
            const int64_t d = deg_of(g, i, v);
            for (int64_t t = 0; t < d; ++t) c->degs[(int64_t)nb[t] * L + i]++;
        }
}

static int cmp_desc_double(const void* a, const void* b) {
    const double x = *(const double*)a, y = *(const double*)b;
    return (x < y) - (x > y);
}

/* FCPathLevelLeft::constructCore wds block (Core/FCPathLevelLeft.cpp:57-99)
 * for the cell whose alive vertices are core[new_e..n). */
static void tree_score(tree_ctx* c, uint32_t k, uint32_t lmd, uint32_t new_e) {
    const orc_graph* g = c->g;
    const int L = g->L;
    const uint32_t n = (uint32_t)g->n;
    const uint32_t length = n - new_e;
    double deg_l[64];
    for (int l = 0; l < L; ++l) {
        uint32_t num_edge = 0; /* uint32 accumulation, wraps like the reference */
        for (uint32_t i = new_e; i < n; ++i) num_edge += (uint32_t)c->degs[(int64_t)c->core[i] * L + l];
        deg_l[l] = (double)((float)num_edge * 1.0f / 2); /* :66 */
    }
    qsort(deg_l, (size_t)L, sizeof(double), cmp_desc_double); /* :72 */
    double layer_density = 0.0;
    int sel = 0;
    for (int l = 0; l < L; ++l) {
        const double mad = (deg_l[l] * 1.0f) / (double)length * 1.0; /* :81 */
        const double dens = mad * c->pow_tab[l];                   /* :82 */
        if (dens >= layer_density) {                               /* :85 */
            layer_density = dens;
            sel = l + 1;
        }
    }
    if (c->cell_density && (int32_t)k < c->cap) c->cell_density[(size_t)(lmd - 1) * c->cap + k] = layer_density;
    if (c->cell_sel && (int32_t)k < c->cap) c->cell_sel[(size_t)(lmd - 1) * c->cap + k] = sel;
    /* FCTree::WeightDenestSubgraph (Core/FCTree.cpp:69-82), float max. */
    if (layer_density >= (double)c->max_density) {
        c->max_density = (float)layer_density;
        c->res_k = k;
        c->res_lmd = lmd;
        c->res_len = length;
    }
}

static void tree_build(tree_ctx* c, uint32_t k, uint32_t lmd, uint32_t e) {
    const orc_graph* g = c->g;
    const int L = g->L;
    const uint32_t n = (uint32_t)g->n;
    const uint32_t s = e, old_e = e;
    uint32_t new_e;
    /* initial violators (FCTreeBuilderLeft.cpp:217-240) */
    for (uint32_t i = s; i < n; ++i) {
        const uint32_t v = c->core[i];
        uint32_t cnt = 0;
        int ok = 0;
        for (int j = 0; j < L; ++j) {
            if (c->degs[(int64_t)v * L + j] >= (int32_t)k) {
                if (++cnt >= lmd) { ok = 1; break; }
            }
        }
        if (!ok) {
            const uint32_t pv = c->pos[v];
            c->core[pv] = c->core[e];
            c->core[e] = v;
            c->pos[c->core[pv]] = pv;
            c->pos[v] = e++;
        }
    }
    new_e = (s != e) ? tree_peel(c, k, lmd, s, e) : e;
    if (n - new_e > 0) {
        c->count++;
        int32_t* row = c->core_index + (size_t)(lmd - 1) * (size_t)n;
        if (lmd <= (uint32_t)L) {
            for (uint32_t i = new_e; i < n; ++i) {
                const uint32_t v = c->core[i];
                if (row[v] < (int32_t)k) row[v] = (int32_t)k;
            }
            if (c->cell_len && (int32_t)k < c->cap) c->cell_len[(size_t)(lmd - 1) * c->cap + k] = n - new_e;
            if (c->kmax[lmd - 1] < (int32_t)k) c->kmax[lmd - 1] = (int32_t)k;
        }
        /* PrintCoreInfor (FCTreeBuilderLeft.cpp:89-112): at lmd = 1 the sum
         * over all vertices of the residual degrees, halved. */
        if (!c->right_spine && lmd == 1) {
            long long s_fc = 0;
            for (int64_t v = 0; v < (int64_t)n; ++v)
                for (int l = 0; l < L; ++l) s_fc += c->degs[v * L + l];
            if (c->edge_fc_k1 && (int32_t)k < c->edge_fc_cap) c->edge_fc_k1[k] = s_fc / 2;
            c->sum_num_of_edge_fc += s_fc / 2;
        }
        if (c->wds) tree_score(c, k, lmd, new_e);
        if (!c->right_spine) {
            tree_build(c, k + 1, lmd, new_e);       /* left (k+1, lmd)  :258-267 */
            if (k == 1) tree_build(c, 1, lmd + 1, new_e); /* right (1, lmd+1) :269-279 */
        } else {
            tree_build(c, k, lmd + 1, new_e);       /* right (k, lmd+1) FCTreeBuilderRight.cpp:220-228 */
            if (lmd == 1) tree_build(c, k + 1, 1, new_e); /* left (k+1, 1)  :230-239 */
        }
    }
    tree_restore(c, old_e, new_e);
}

int orc_tree(int64_t n, int32_t L, const int64_t* const* row_ptr, const uint32_t* const* col, int32_t right_spine,
             int32_t* core_index, int64_t* cell_len, int32_t cap, int32_t* kmax, int64_t* count,
             int64_t* sum_num_of_edge_fc, int64_t* edge_fc_k1, int32_t edge_fc_cap) {
    if (n <= 0 || L <= 0 || L > 64) return -1;
    orc_graph g = {n, L, row_ptr, col};
    tree_ctx c;
    memset(&c, 0, sizeof c);
    c.g = &g;
    c.degs = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n * L));
    c.core = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    c.pos = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    if (!c.degs || !c.core || !c.pos) { free(c.degs); free(c.core); free(c.pos); return -3; }
    for (int64_t v = 0; v < n; ++v) {
        for (int l = 0; l < L; ++l) c.degs[v * L + l] = (int32_t)deg_of(&g, l, v);
        c.core[v] = (uint32_t)v;
        c.pos[v] = (uint32_t)v;
    }
    memset(core_index, 0, sizeof(int32_t) * (size_t)(n * L));
    if (cell_len) {
        memset(cell_len, 0, sizeof(int64_t) * (size_t)L * (size_t)cap);
        if (cap > 0) for (int l = 0; l < L; ++l) cell_len[(size_t)l * cap] = n;
    }
    for (int l = 0; l < L; ++l) kmax[l] = 0;
    c.core_index = core_index;
    c.cell_len = cell_len;
    c.cap = cap;
    c.kmax = kmax;
    c.right_spine = right_spine;
    c.edge_fc_k1 = edge_fc_k1;
    c.edge_fc_cap = edge_fc_cap;
    tree_build(&c, 1, 1, 0);
    if (count) *count = c.count;
    if (sum_num_of_edge_fc) *sum_num_of_edge_fc = c.sum_num_of_edge_fc;
    free(c.degs); free(c.core); free(c.pos);
    return 0;
}

/* WdsList result (FCPathLevelLeft::Execute with wds = true, then
 * FCTree::WeightDenestSubgraph, main.cpp:402-424), computed on the lpe DFS,
 * which visits the same cells with the same residual degrees in the same
 * lambda-major, k-ascending order.  pow_tab[l] = pow(l+1, beta) as the
 * reference computes it (FCPathLevelLeft.cpp:82). */
int orc_wds(int64_t n, int32_t L, const int64_t* const* row_ptr, const uint32_t* const* col, const double* pow_tab,
            float* max_density, uint32_t* res_k, uint32_t* res_lmd, uint32_t* res_len,
            double* cell_density, int32_t* cell_sel, int32_t cap) {
    if (n <= 0 || L <= 0 || L > 64) return -1;
    orc_graph g = {n, L, row_ptr, col};
    tree_ctx c;
    memset(&c, 0, sizeof c);
    c.g = &g;
    c.degs = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n * L));
    c.core = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    c.pos = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    int32_t* ci = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n * L));
    int32_t* km = (int32_t*)malloc(sizeof(int32_t) * (size_t)L);
    if (!c.degs || !c.core || !c.pos || !ci || !km) {
        free(c.degs); free(c.core); free(c.pos); free(ci); free(km);
        return -3;
    }
    for (int64_t v = 0; v < n; ++v) {
        for (int l = 0; l < L; ++l) c.degs[v * L + l] = (int32_t)deg_of(&g, l, v);
        c.core[v] = (uint32_t)v;
        c.pos[v] = (uint32_t)v;
    }
    memset(ci, 0, sizeof(int32_t) * (size_t)(n * L));
    for (int l = 0; l < L; ++l) km[l] = 0;
    if (cell_density) memset(cell_density, 0, sizeof(double) * (size_t)L * (size_t)cap);
    if (cell_sel) memset(cell_sel, 0, sizeof(int32_t) * (size_t)L * (size_t)cap);
    c.core_index = ci;
    c.kmax = km;
    c.wds = 1;
    c.pow_tab = pow_tab;
    c.max_density = 0.0f;
    c.cell_density = cell_density;
    c.cell_sel = cell_sel;
    c.cap = cap;
    tree_build(&c, 1, 1, 0);
    *max_density = c.max_density;
    *res_k = c.res_k;
    *res_lmd = c.res_lmd;
    *res_len = c.res_len;
    free(c.degs); free(c.core); free(c.pos); free(ci); free(km);
    return 0;
}

/* --------------------------------------------------------------- CoreIndex
 * Restates get_index2 + findLmdthLargest (CoreIndexParallel/CoreIndex.cpp:
 * 46-109) for one lmd.  CoreIndex::Execute (:168-213) runs it for lmd =
 * 1..L one after another on fresh degree copies; CoreIndex::ExecuteParallel
 * (:217-300, "ParCoreIndex") runs the lmds as parallel loop iterations, each
 * on its own copy -- orc_coreindex_lmd is one such iteration and may be
 * called from several threads at once (it shares nothing but the read-only
 * graph).  Bucket peeling on the lmd-th largest layer degree; std::set
 * buckets become bucket stacks with lazy deletion (a vertex's live bucket is
 * bucket_of[v]); the index does not depend on the pop order within a bucket.
 * The neighbour loop decrements every neighbour's layer degree and masks the
 * step for neighbours already at or below k (the reference skips them with
 * `continue`): their degrees are never read again, so the result is the
 * same, and the loop has no data-dependent branch on a cache miss.  Large
 * arrays are requested in transparent huge pages (a 100M-vertex graph's
 * degree array is 3.2 GB of random accesses). */
static uint32_t lmd_th_largest(const uint32_t* deg, uint32_t lmd, int L, uint32_t* tmp) {
    memcpy(tmp, deg, sizeof(uint32_t) * (size_t)L);
    /* partial selection sort, descending, first lmd entries */
    for (uint32_t i = 0; i < lmd; ++i) {
        int best = (int)i;
        for (int j = (int)i + 1; j < L; ++j)
            if (tmp[j] > tmp[best]) best = j;
        const uint32_t t = tmp[i];
        tmp[i] = tmp[best];
        tmp[best] = t;
    }
    return tmp[lmd - 1];
}

typedef struct {
    uint32_t* v;
    int64_t size, cap;
} ustack;

static int ustack_push(ustack* s, uint32_t x) {
    if (s->size == s->cap) {
        int64_t nc = s->cap ? s->cap * 2 : 4;
        uint32_t* nv = (uint32_t*)realloc(s->v, sizeof(uint32_t) * (size_t)nc);
        if (!nv) return -1;
        s->v = nv;
        s->cap = nc;
    }
    s->v[s->size++] = x;
    return 0;
}

static void* big_alloc(size_t bytes) {
    if (bytes == 0) bytes = 1;
    void* p = mmap(NULL, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (p == MAP_FAILED) return NULL;
#ifdef MADV_HUGEPAGE
    madvise(p, bytes, MADV_HUGEPAGE);
#endif
    return p;
}

static void big_free(void* p, size_t bytes) {
    if (p) munmap(p, bytes ? bytes : 1);
}

/* Vertex-major merged adjacency for the CoreIndex loop: the entries of v in
 * every layer are adj[off[v] .. off[v+1]), layer 0 first, each entry stored
 * as u * L + l -- the index of degs[u][l] in a row-major [n][L] array.  A
 * popped vertex's neighbourhood is then one contiguous run instead of L
 * separate rows (the reference walks adj_lst[vv] of every layer,
 * CoreIndex.cpp:84-95; same entries, same order within a layer).  Requires
 * n * L < 2^32.  off has n+1 entries, adj sum of all row lengths. */
int orc_merge(int64_t n, int32_t L, const int64_t* const* row_ptr, const uint32_t* const* col, int64_t* off,
              uint32_t* adj) {
    if (n <= 0 || L <= 0 || L > 64 || (uint64_t)n * (uint64_t)L >= (1ull << 32)) return -1;
    int64_t e = 0;
    for (int64_t v = 0; v < n; ++v) {
        off[v] = e;
        for (int l = 0; l < L; ++l) {
            const uint32_t* nb = col[l] + row_ptr[l][v];
            const int64_t d = row_ptr[l][v + 1] - row_ptr[l][v];
            for (int64_t t = 0; t < d; ++t) adj[e++] = nb[t] * (uint32_t)L + (uint32_t)l;
        }
    }
    off[n] = e;
    return 0;
}

/* get_index2 for one lmd over the merged adjacency (see above). */
int orc_coreindex_merged_lmd(int64_t n, int32_t L, const int64_t* off, const uint32_t* adj, int32_t lmd,
                             int32_t* core_row) {
    if (n <= 0 || L <= 0 || L > 64 || lmd < 1 || lmd > L) return -1;
    const size_t degs_b = sizeof(uint32_t) * (size_t)n * (size_t)L;
    uint32_t* degs = (uint32_t*)big_alloc(degs_b);
    uint32_t* iv = (uint32_t*)big_alloc(sizeof(uint32_t) * (size_t)n);
    int32_t* bucket_of = (int32_t*)big_alloc(sizeof(int32_t) * (size_t)n);
    uint8_t* in_n = (uint8_t*)big_alloc((size_t)n);
    uint32_t* nlist = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    ustack* B = NULL;
    uint32_t tmp[64];
    uint32_t kcap = 0;
    const uint32_t uL = (uint32_t)L;
    int rc = 0;
    if (!degs || !iv || !bucket_of || !in_n || !nlist) { rc = -3; goto done; }
    /* the degree copy of CoreIndex.cpp:181-191 / :254-258 */
    memset(degs, 0, degs_b);
    for (int64_t v = 0; v < n; ++v)
        for (int64_t t = off[v]; t < off[v + 1]; ++t) degs[v * L + (adj[t] % uL)] += 1;
    /* CoreIndex.cpp:71-74 */
    for (int64_t v = 0; v < n; ++v) {
        iv[v] = lmd_th_largest(degs + v * L, (uint32_t)lmd, L, tmp);
        if (iv[v] > kcap) kcap = iv[v];
    }
    B = (ustack*)calloc((size_t)kcap + 1, sizeof(ustack));
    if (!B) { rc = -3; goto done; }
    for (int64_t v = 0; v < n; ++v) {
        bucket_of[v] = (int32_t)iv[v];
        if (ustack_push(&B[iv[v]], (uint32_t)v)) { rc = -3; goto done; }
    }
    /* CoreIndex.cpp:77-104; buckets above kcap stay empty */
    for (int64_t k = 0; k <= (int64_t)kcap; ++k) {
        while (B[k].size > 0) {
            const uint32_t vv = B[k].v[--B[k].size];
            if (bucket_of[vv] != k) continue; /* stale entry */
            bucket_of[vv] = -1;               /* popped */
            iv[vv] = (uint32_t)k;
            int64_t nn = 0;
            for (int64_t t = off[vv]; t < off[vv + 1]; ++t) {
                const uint32_t e = adj[t];
                const uint32_t u = e / uL;
                const uint32_t ivu = iv[u];
                const uint32_t live = ivu > (uint32_t)k; /* i_v[nb_u] <= k: continue */
                const uint32_t nd = degs[e] - live;
                degs[e] = nd;
                if ((live & (nd + 1 == ivu)) && !in_n[u]) {
                    in_n[u] = 1;
                    nlist[nn++] = u;
                }
            }
            for (int64_t t = 0; t < nn; ++t) {
                const uint32_t u = nlist[t];
                in_n[u] = 0;
                iv[u] = lmd_th_largest(degs + (int64_t)u * L, (uint32_t)lmd, L, tmp);
                const int64_t nb = iv[u] > (uint32_t)k ? (int64_t)iv[u] : k;
                bucket_of[u] = (int32_t)nb;
                if (ustack_push(&B[nb], u)) { rc = -3; goto done; }
            }
        }
    }
    for (int64_t v = 0; v < n; ++v) core_row[v] = (int32_t)iv[v];
done:
    if (B) {
        for (int64_t b = 0; b <= (int64_t)kcap; ++b) free(B[b].v);
        free(B);
    }
    big_free(degs, degs_b);
    big_free(iv, sizeof(uint32_t) * (size_t)n);
    big_free(bucket_of, sizeof(int32_t) * (size_t)n);
    big_free(in_n, (size_t)n);
    free(nlist);
    return rc;
}

static int64_t total_entries(int64_t n, int32_t L, const int64_t* const* row_ptr) {
    int64_t m = 0;
    for (int l = 0; l < L; ++l) m += row_ptr[l][n] - row_ptr[l][0];
    return m;
}

/* CoreIndex::Execute for the listed lmds (lmds == NULL: 1..L); the row of
 * lmds[i] goes to core_index + i * n. */
static int coreindex_lmds(int64_t n, int32_t L, const int64_t* const* row_ptr, const uint32_t* const* col,
                          const int32_t* lmds, int32_t count, int32_t* core_index) {
    if (n <= 0 || L <= 0 || L > 64 || (uint64_t)n * (uint64_t)L >= (1ull << 32)) return -1;
    const int64_t m = total_entries(n, L, row_ptr);
    int64_t* off = (int64_t*)big_alloc(sizeof(int64_t) * (size_t)(n + 1));
    uint32_t* adj = (uint32_t*)big_alloc(sizeof(uint32_t) * (size_t)m);
    int rc = (!off || !adj) ? -3 : orc_merge(n, L, row_ptr, col, off, adj);
    for (int32_t i = 0; rc == 0 && i < count; ++i) {
        const int32_t lmd = lmds ? lmds[i] : i + 1;
        rc = orc_coreindex_merged_lmd(n, L, off, adj, lmd, core_index + (size_t)i * (size_t)n);
    }
    big_free(off, sizeof(int64_t) * (size_t)(n + 1));
    big_free(adj, sizeof(uint32_t) * (size_t)m);
    return rc;
}

/* One lmd of CoreIndex (one iteration of ParCoreIndex's loop) into core_row. */
int orc_coreindex_lmd(int64_t n, int32_t L, const int64_t* const* row_ptr, const uint32_t* const* col, int32_t lmd,
                      int32_t* core_row) {
    if (lmd < 1 || lmd > L) return -1;
    return coreindex_lmds(n, L, row_ptr, col, &lmd, 1, core_row);
}

/* CoreIndex::Execute (CoreIndex.cpp:168-213): every lmd in turn. */
int orc_coreindex(int64_t n, int32_t L, const int64_t* const* row_ptr, const uint32_t* const* col,
                  int32_t* core_index) {
    return coreindex_lmds(n, L, row_ptr, col, NULL, L, core_index);
}

/* ------------------------------------------------------ local FirmCore check
 * Two local conditions every FirmCore index row c (one lmd) satisfies,
 * independent of any peeling routine; counts the vertices of [v0, v1) that
 * violate one of them:
 *   (valid) every v with c(v) = k >= 1 has at least lmd layers l with
 *           |{u in N_l(v) : c(u) >= k}| >= k -- S(k, lmd) = {c >= k} meets
 *           the FirmCore condition at v;
 *   (tight) every v with c(v) = k has fewer than lmd layers l with
 *           |{u in N_l(v) : c(u) >= k+1}| >= k+1 -- else S(k+1, lmd) plus v
 *           would be a larger (k+1, lmd) set.
 * (valid) alone is met by a row that is too low everywhere and (tight) alone
 * by one that is too high, so together they reject an off-by-one in either
 * direction at any vertex.  first_bad (may be NULL) gets the smallest
 * violating id, or -1.  Read-only; threads may split [0, n). */
int64_t orc_check_lmd(int64_t n, int32_t L, const int64_t* const* row_ptr, const uint32_t* const* col, int32_t lmd,
                      const int32_t* c, int64_t v0, int64_t v1, int64_t* first_bad) {
    if (first_bad) *first_bad = -1;
    if (n <= 0 || L <= 0 || lmd < 1 || lmd > L || v0 < 0 || v1 > n) return -1;
    int64_t bad = 0;
    for (int64_t v = v0; v < v1; ++v) {
        const int32_t k = c[v];
        int good = 0, over = 0;
        for (int l = 0; l < L; ++l) {
            const uint32_t* nb = col[l] + row_ptr[l][v];
            const int64_t d = row_ptr[l][v + 1] - row_ptr[l][v];
            int64_t ge = 0, gt = 0;
            for (int64_t t = 0; t < d; ++t) {
                const int32_t cu = c[nb[t]];
                ge += cu >= k;
                gt += cu > k;
            }
            good += ge >= k;
            over += gt >= (int64_t)k + 1;
        }
        if ((k >= 1 && good < lmd) || over >= lmd || k < 0) {
            if (first_bad && *first_bad < 0) *first_bad = v;
            ++bad;
        }
    }
    return bad;
}

/* ------------------------------------------------------------ WdsCoreIndex
 * Per lmd, the result get_wds leaves in its MaxDensity record
 * (CoreIndexParallel/CoreIndex.cpp:347-413): for k = 1, 2, ... while
 * S(k, lmd) = {v : index >= k} is not empty, calculateWeightDenestSubgraph-
 * Common (:307-345) scores the cell in float: per layer a float counter
 * incremented once per directed entry inside S, halved; sorted descending;
 * density_l = (float)((degree_l / |S|) * pow(l+1, beta)) with the float
 * beta promoted to double; the record takes (density, k, lmd) on >=.
 * The record starts at {0, 0, 0.0f}.  Indices come from orc_coreindex. */
static int cmp_desc_float(const void* a, const void* b) {
    const float x = *(const float*)a, y = *(const float*)b;
    return x < y ? 1 : (x > y ? -1 : 0);
}

int orc_wds_coreindex(int64_t n, int32_t L, const int64_t* const* row_ptr, const uint32_t* const* col, float beta,
                      float* density, uint32_t* res_k, uint32_t* res_lmd) {
    if (n <= 0 || L <= 0 || L > 64) return -1;
    orc_graph g = {n, L, row_ptr, col};
    int32_t* ci = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n * L));
    if (!ci) return -3;
    int rc = orc_coreindex(n, L, row_ptr, col, ci);
    if (rc) {
        free(ci);
        return rc;
    }
    for (int32_t lmd = 1; lmd <= L; ++lmd) {
        const int32_t* idx = ci + (size_t)(lmd - 1) * (size_t)n;
        float best = 0.0f;
        uint32_t bk = 0, bl = 0;
        int32_t kmax = 0;
        for (int64_t v = 0; v < n; ++v)
            if (idx[v] > kmax) kmax = idx[v];
        for (int32_t k = 1; k <= kmax; ++k) {
            uint32_t n_node = 0;
            for (int64_t v = 0; v < n; ++v) n_node += idx[v] >= k;
            float deg[64];
            for (int l = 0; l < L; ++l) {
                float n_edge = 0;
                for (int64_t v = 0; v < n; ++v) {
                    if (idx[v] < k) continue;
                    const uint32_t* nb = g.col[l] + g.row_ptr[l][v];
                    const int64_t d = deg_of(&g, l, v);
                    for (int64_t t = 0; t < d; ++t)
                        if (idx[nb[t]] >= k) n_edge++;
                }
                deg[l] = n_edge / 2;
            }
            qsort(deg, (size_t)L, sizeof(float), cmp_desc_float);
            for (int l = 0; l < L; ++l) {
                const float mad = deg[l] / n_node;
                const float dens = (float)(mad * pow((double)(l + 1), (double)beta));
                if (dens >= best) {
                    best = dens;
                    bk = (uint32_t)k;
                    bl = (uint32_t)lmd;
                }
            }
        }
        density[lmd - 1] = best;
        res_k[lmd - 1] = bk;
        res_lmd[lmd - 1] = bl;
    }
    free(ci);
    return 0;
}

/* ------------------------------------------------------------- brute force
 * Definition-level fixed point, independent of every reference routine:
 * for each (k, lmd), start from all vertices and repeatedly delete every v
 * with fewer than lmd layers l where |N_l(v) ∩ S| >= k, until nothing
 * changes.  S(k, lmd) is what remains; index = largest such k.  Small graphs
 * only (cost ~ kmax * rounds * edges). */
int orc_bruteforce(int64_t n, int32_t L, const int64_t* const* row_ptr, const uint32_t* const* col,
                   int32_t* core_index) {
    if (n <= 0 || L <= 0) return -1;
    orc_graph g = {n, L, row_ptr, col};
    uint8_t* in_s = (uint8_t*)malloc((size_t)n);
    uint8_t* nxt = (uint8_t*)malloc((size_t)n);
    if (!in_s || !nxt) { free(in_s); free(nxt); return -3; }
    memset(core_index, 0, sizeof(int32_t) * (size_t)(n * L));
    for (int32_t lmd = 1; lmd <= L; ++lmd) {
        int32_t* row = core_index + (size_t)(lmd - 1) * (size_t)n;
        for (int64_t k = 1;; ++k) {
            memset(in_s, 1, (size_t)n);
            int changed = 1;
            while (changed) {
                changed = 0;
                for (int64_t v = 0; v < n; ++v) {
                    nxt[v] = in_s[v];
                    if (!in_s[v]) continue;
                    int good = 0;
                    for (int l = 0; l < L; ++l) {
                        int64_t d = 0;
                        for (int64_t t = g.row_ptr[l][v]; t < g.row_ptr[l][v + 1]; ++t) d += in_s[g.col[l][t]];
                        good += d >= k;
                    }
                    if (good < lmd) { nxt[v] = 0; changed = 1; }
                }
                memcpy(in_s, nxt, (size_t)n);
            }
            int64_t any = 0;
            for (int64_t v = 0; v < n; ++v)
                if (in_s[v]) { row[v] = (int32_t)k; ++any; }
            if (!any) break;
        }
    }
    free(in_s); free(nxt);
    return 0;
}
