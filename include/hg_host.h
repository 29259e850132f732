/*
 * hg_host.h -- host-side graph formats for the HipGrid FirmCore engine.
 *
 * Plain C ABI (no C++ or HIP types).  Implemented in
 * pargridcpu_amd/csrc/hg_host.cpp, built into libhg_host.so.
 *
 * A multilayer graph here is L layers over one vertex set 0..n-1.  Each layer
 * is a symmetric CSR: row_ptr[l] has n+1 int64 entries, col[l] holds the
 * neighbour ids of every vertex in ascending order, de-duplicated, without
 * self-loops, each undirected edge stored in both directions.  This is the
 * reference's degree-prefixed row layout (Graph/Graph.cpp:17-50, row v =
 * [deg, nbrs...]) split into offsets + ids: deg_l(v) = row_ptr[l][v+1] -
 * row_ptr[l][v] and nbrs = col[l][row_ptr[l][v] .. row_ptr[l][v+1]).
 */
#ifndef HG_HOST_H
#define HG_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes shared with hipgrid.h. */
#define HGH_OK 0
#define HGH_ERR_ARG (-1)
#define HGH_ERR_IO (-2)
#define HGH_ERR_NOMEM (-3)
#define HGH_ERR_RANGE (-4)

typedef struct hgh_graph hgh_graph;

/* Seeded Chung-Lu multilayer generator parameters.
 * No counterpart in the reference (it only reads edge-list files); this is
 * the synthetic input the benchmark configurations are defined on.
 *   weights   w_i = (i+1)^(-1/(gamma-1)), i = 0..n-1
 *   reshuffle floor(rho*n) positions chosen by a partial Fisher-Yates pass,
 *             their weights permuted among themselves, independently per layer
 *   cap       if cap_deg > 0: w_i <- min(w_i, cap_deg * sum(w) / (2m)) (one pass)
 *   edges     m endpoint pairs drawn i.i.d. proportional to w (Vose alias);
 *             self-loops dropped, duplicates merged, stored symmetric.
 *   m         m_lo == m_hi: m = m_lo for every layer; otherwise m is
 *             log-uniform in [m_lo, m_hi], drawn per layer.
 * PRNG: SplitMix64, one counter-based substream per (layer, purpose), so the
 * output does not depend on the thread count. */
typedef struct hgh_gen_params {
    uint64_t seed;
    int64_t n;
    int32_t L;
    double gamma;
    double rho;
    int64_t m_lo;
    int64_t m_hi;
    double cap_deg;
} hgh_gen_params;

/* Generates a graph; *out is owned by the caller (hgh_graph_free). */
int hgh_generate(const hgh_gen_params* p, hgh_graph** out);

/* Loads the reference's on-disk format: <dir>/mlg.conf lists one layer file
 * per line; each layer file holds whitespace-separated "u v" pairs.  Vertex
 * ids are assigned in first-appearance order over the layers in conf order;
 * self-loops are dropped; duplicates merged.  Replaces
 * MultilayerGraph::LoadFromFile / GetGraphFile / LoadLayer
 * (Graph/MultilayerGraph.cpp:47-166) and Graph::BuildFromEdgeLst
 * (Graph/Graph.cpp:17-50).  If write_vtx_map != 0, writes
 * <dir>/__vtx_map.txt ("label id" per line, first-appearance order) as the
 * reference does (MultilayerGraph.cpp:70, 133-147). */
int hgh_load_text(const char* dir, int32_t write_vtx_map, hgh_graph** out);

/* Writes the reference's on-disk format for g into dir (which must exist):
 * mlg.conf naming layer_<l>.txt, each listing every undirected edge once as
 * "u v" with u < v, using the vertex labels of g (see hgh_graph_label). */
int hgh_write_text(const hgh_graph* g, const char* dir);

/* Layer order as MultilayerGraph::SetGraphOrder (MultilayerGraph.cpp:9-45):
 * o = 0 keeps file order; 1/2 sort by increasing/decreasing float density
 * (m_l/2)/n; 3/4 by increasing/decreasing directed entry count m_l.  Fills
 * order[0..L) with original layer numbers; layer i of the ordered view is
 * original layer order[i]. */
int hgh_layer_order(const hgh_graph* g, int32_t o, int32_t* order);

/* Returns a new graph whose layer i is layer order[i] of g. */
int hgh_permute_layers(const hgh_graph* g, const int32_t* order, hgh_graph** out);

/* Accessors.  Pointers stay valid until hgh_graph_free. */
int64_t hgh_graph_n(const hgh_graph* g);
int32_t hgh_graph_layers(const hgh_graph* g);
const int64_t* hgh_graph_row_ptr(const hgh_graph* g, int32_t layer);
const uint32_t* hgh_graph_col(const hgh_graph* g, int32_t layer);
int64_t hgh_graph_entries(const hgh_graph* g, int32_t layer); /* directed entries m_l */
/* Original vertex label of id v (identity for generated graphs; the file's
 * label for loaded ones, like the reference's id2vtx map, main.cpp:134-135). */
uint64_t hgh_graph_label(const hgh_graph* g, int64_t v);
void hgh_graph_free(hgh_graph* g);

/* "|L| = .., |V| = .., |E| = .., max_|E_i| = .." statistics line as
 * MultilayerGraph::PrintStatistics (MultilayerGraph.cpp:168-178), with the
 * reference's 32-bit |E| sums.  Writes at most cap bytes into buf. */
int hgh_statistics(const hgh_graph* g, char* buf, int32_t cap);

/* Peak resident set size in MB, as GetPeakRSSInMB (Util/MemoryUtils.h:44-124). */
double hgh_peak_rss_mb(void);

const char* hgh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HG_HOST_H */
