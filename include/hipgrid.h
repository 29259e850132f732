/*
 * hipgrid.h -- C ABI of the HipGrid FirmCore engine (libhipgrid.so).
 *
 * Replaces the reference's OpenMP decomposition entry points for one GPU:
 *   FCSyncLeft::Execute(MultilayerGraph&, FCCoreTree&)      (-m ParGrid,
 *       CoreParallel/FCSyncLeft.h:18, FCSyncLeft.cpp:202-252, main.cpp:229-257)
 *   FCSyncLeft::ExecuteMix / FCSyncRight::Execute(Mix) / FCSyncMem::Execute
 *       (same results: FCSyncLeft.h:34, FCSyncRight.h:17,33, FCSyncMem.h:17)
 *   FCPathLevelLeft::Execute(mg, tree, wds=true, beta) + FCTree::WeightDenestSubgraph
 *       (-m WdsList, Core/FCPathLevelLeft.cpp:311-377, Core/FCTree.cpp:69-82,
 *        main.cpp:402-442)
 * with plain pointers and sizes: no C++, HIP or torch types in any signature.
 *
 * Results (the reference's fcgrid tree, flattened):
 *   core_index[lmd-1][v] = max{k >= 1 : v in S(k, lmd)}, 0 if v is not in
 *       S(1, lmd), for lmd = 1..L, where S(k, lmd) is the (k, lmd)-FirmCore:
 *       the largest vertex set in which every vertex has >= k neighbours
 *       inside the set in at least lmd layers.  int32, lambda-major.
 *   cell length |S(k, lmd)| = #{v : core_index[lmd-1][v] >= k}  (coreNodeP::
 *       length, FCCoreTree.h:10-12); the non-empty cells are exactly
 *       1 <= k <= kmax[lmd-1].
 * Every result is bit-exact against the reference on the same input.
 *
 * Errors: every call returns 0 on success and a negative HG_ERR_* code on
 * failure (the reference has no error channel: it aborts on OOM and ignores
 * unknown methods); hg_last_error() returns the message of the calling
 * thread's last failure.  No call aborts the process.
 *
 * Threading: a handle owns its device memory and uses the HIP stream given
 * at creation (NULL: the device's default stream).  Distinct handles may be
 * used from distinct threads; one handle must not be used by two threads at
 * once.  All calls that return results synchronise with the stream.
 */
#ifndef HIPGRID_H
#define HIPGRID_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HG_OK 0
#define HG_ERR_ARG (-1)     /* bad argument */
#define HG_ERR_HIP (-2)     /* HIP runtime error */
#define HG_ERR_NOMEM (-3)   /* device or host allocation failed */
#define HG_ERR_RANGE (-4)   /* size limit exceeded (see hg_graph_create) */
#define HG_ERR_STATE (-5)   /* call out of order (e.g. results before decompose) */
#define HG_ERR_DEVICE (-6)  /* device-side check failed (watchdog, overflow) */

/* One layer as symmetric CSR: row_ptr has n+1 entries (int32 or int64 as
 * chosen by offset_bits), col[row_ptr[v] .. row_ptr[v+1]) lists the
 * neighbours of v, each undirected edge stored in both directions, no
 * self-loops, no duplicates.  This is the reference's degree-prefixed row
 * adj_lst[v] = [deg, nbrs...] (Graph/Graph.cpp:17-50, MultilayerGraph.h:38-40)
 * split into offsets and ids.  Host memory. */
typedef struct hg_layer_csr {
    const void* row_ptr;
    const uint32_t* col;
} hg_layer_csr;

typedef struct hg_graph hg_graph;

/* Uploads the graph (layer i of `layers` is the reference's GetGraph(i),
 * i.e. layers already in -o order) and allocates the working memory for all
 * lmd = 1..L.  offset_bits is 32 or 64.  Limits: 1 <= L <= 64,
 * 1 <= n, n * L < 2^32, total directed entries < 2^40.  stream is a
 * hipStream_t (NULL: default stream).  Replaces the MultilayerGraph
 * hand-over into FCSyncLeft::Execute (FCSyncLeft.cpp:202-235: degs init and
 * valid init happen here and in hg_decompose). */
int hg_graph_create(const hg_layer_csr* layers, int32_t L, int64_t n, int32_t offset_bits, void* stream,
                    hg_graph** out);

/* Full decomposition for lmd = 1..L on the device; results stay resident in
 * device memory until the next hg_decompose or hg_graph_destroy.  Returns
 * after the device work has finished.  Replaces FCSyncLeft::Execute
 * (FCSyncLeft.cpp:202-252) and its siblings listed above. */
int hg_decompose(hg_graph* g);

/* Copies core_index[L][n] (int32, lambda-major) to host memory. */
int hg_core_index(hg_graph* g, int32_t* core_index);

/* Device pointer to the resident core_index[L][n] (valid until the next
 * hg_decompose or hg_graph_destroy). */
int hg_core_index_device(hg_graph* g, const int32_t** dev_ptr);

/* Cell lengths: cell_len[(lmd-1)*cap + k] = |S(k, lmd)| for k < cap
 * (k = 0 gives n), kmax[lmd-1] = largest non-empty k (0 if none).  Either
 * output may be NULL.  Replaces the per-node length/total fields of the
 * fcgrid tree (constructCoreSync, FCSyncLeft.cpp:10-38). */
int hg_cells(hg_graph* g, int64_t* cell_len, int32_t cap, int32_t* kmax);

/* k-shells as FCSyncMem keeps them (father->diff_res = S(k,lmd) \ S(k+1,lmd),
 * CoreParallel/FCSyncMem.cpp:42-51): shell_off has L*(kcap+1)+1 entries;
 * the vertices with core_index[lmd-1][v] == k are
 * shell_vtx[shell_off[(lmd-1)*(kcap+1) + k] .. shell_off[... + k + 1]),
 * ascending.  shell_vtx must hold n*L entries.  kcap must exceed every kmax. */
int hg_shells(hg_graph* g, int32_t kcap, int64_t* shell_off, uint32_t* shell_vtx);

/* Weighted densest subgraph over the fcgrid cells, as -m WdsList computes
 * it (Core/FCPathLevelLeft.cpp:57-99 per cell, Core/FCTree.cpp:69-82 and
 * main.cpp:402-424 for the argmax): per cell S = S(k, lmd) and layer l,
 * num_edge_l = sum over v in S of |N_l(v) ∩ S| as uint32; degree_l =
 * (float)num_edge_l / 2; sorted descending; density(cell) = max over i of
 * degree_(i)/|S| * pow(i+1, beta) (double, >=); the answer is the cell with
 * the largest density under a float running max updated with >= in
 * lambda-major, k-ascending order.  Requires hg_decompose first. */
typedef struct hg_wds {
    float max_density;       /* maximum_density (main.cpp:406) */
    uint32_t res_k;          /* res_k */
    uint32_t res_lmd;        /* res_lmd */
    uint32_t res_len;        /* res_len = |S(res_k, res_lmd)| */
    int32_t num_select_layer; /* coreNode::num_select_layer of that cell */
    int32_t reserved;
} hg_wds;
int hg_densest(hg_graph* g, double beta, hg_wds* out);

/* Per-cell densities of the last hg_densest: density[(lmd-1)*cap + k] for
 * k < cap (0 where the cell is empty or k = 0).  Either output may be NULL. */
int hg_cell_density(hg_graph* g, double* density, int32_t* num_select_layer, int32_t cap);

/* Vertex set S(res_k, res_lmd) of the last hg_densest, ascending ids.  The
 * reference keeps it only as the tree node's core list. */
int hg_densest_vertices(hg_graph* g, uint32_t* out, int64_t cap, int64_t* count);

/* Device time of the phases of the last hg_decompose / hg_densest in ms,
 * measured with HIP events on the handle's stream. */
typedef struct hg_timing {
    double decompose_ms;  /* whole hg_decompose, first to last kernel */
    double init_ms;       /* level-0 scan (degree init) */
    double scan_ms;       /* sum of level scans k >= 1 */
    double cascade_ms;    /* sum of removal cascades */
    double densest_ms;    /* edge histogram + cell scoring of hg_densest */
    int64_t levels;       /* level iterations run */
    int64_t launches;     /* kernel launches in hg_decompose */
} hg_timing;
int hg_last_timing(hg_graph* g, hg_timing* out);

/* Graph facts for reporting. */
int hg_graph_info(hg_graph* g, int64_t* n, int32_t* L, int64_t* directed_entries, int64_t* device_bytes);

int hg_graph_destroy(hg_graph* g);

const char* hg_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HIPGRID_H */
