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
#define HG_ERR_CANCELLED (-7) /* the request was cancelled before it started */

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
 * 1 <= n < 2^31 - 1, n * L < 2^32, total directed entries < 2^40
 * (HG_ERR_RANGE otherwise).  stream is a
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

/* -m WdsCoreIndex (CoreIndex::WdsCoreIndex, CoreIndexParallel/CoreIndex.cpp:
 * 307-478; main.cpp:444-451): the same cells scored with that routine's
 * float arithmetic (float edge counters, which stop at 2^24; float beta;
 * float densities).  per_lmd[lmd-1] (may be NULL) is the best cell of each
 * lmd, k ascending, updated on >=.  *answer (may be NULL) is what the
 * reference prints: the first record with the largest density among
 * records 0..L-1, where record 0 is all zeros (the reference leaves it
 * unwritten) and record lmd is per_lmd[lmd-1] -- so lmd = L never takes
 * part, and for L = 1 the answer is record 0.  res_len is |S(res_k,
 * res_lmd)|; num_select_layer the layer count of the best density.
 * Requires hg_decompose first. */
int hg_densest_coreindex(hg_graph* g, double beta, hg_wds* answer, hg_wds* per_lmd);

/* hg_densest without any device or pinned-memory allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * during the call (code around HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same results as hg_densest, and
 * hg_cell_density / hg_densest_vertices report on it the same way.  The
 * pass's edge histograms, cell lengths, scores and pow table live in one
 * fixed device array of HG_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; a graph needs 8 * L * (kmax + 1) + 28 * (kmax + 1) + 8 * L
 * bytes of it (plus alignment), kmax the largest core index of any lmd, and
 * a larger need fails with HG_ERR_RANGE.  Host copies are plain copies of
 * the caller's memory (no staging buffers).  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (from two threads, on any
 * handles) race and must be serialised by the caller.  Requires
 * hg_decompose first. */
#define HG_NOALLOC_SCRATCH_BYTES (256 * 1024)
int hg_densest_noalloc(hg_graph* g, double beta, hg_wds* out);

/* Per-cell densities of the last hg_densest: density[(lmd-1)*cap + k] for
 * k < cap (0 where the cell is empty or k = 0).  Either output may be NULL. */
int hg_cell_density(hg_graph* g, double* density, int32_t* num_select_layer, int32_t cap);

/* Vertex set S(res_k, res_lmd) of the last hg_densest, ascending ids.  The
 * reference keeps it only as the tree node's core list. */
int hg_densest_vertices(hg_graph* g, uint32_t* out, int64_t cap, int64_t* count);

/* Device time of the phases of the last hg_decompose / hg_densest in ms,
 * measured with HIP events on the handle's stream.  scan_ms and cascade_ms
 * are measured only with phase timing on (hg_set_phase_timing), else -1;
 * the reference prints the analogous split per peel (FCSyncLeft.cpp:
 * 240-245, U_overall). */
typedef struct hg_timing {
    double decompose_ms;  /* whole hg_decompose, first to last kernel */
    double init_ms;       /* step 0: level-0 pass (degree init) and its removal cascade */
    double scan_ms;       /* sum of the scans of steps 1, 2, ...; -1 unless phase timing is on */
    double cascade_ms;    /* sum of the cascades of steps 1, 2, ...; -1 unless phase timing is on */
    double densest_ms;    /* edge histogram + cell scoring of hg_densest */
    int64_t levels;       /* level iterations run */
    int64_t launches;     /* kernel launches in hg_decompose */
    int64_t binned;       /* (instance, round) pairs applied binned by target range in LDS */
} hg_timing;
int hg_last_timing(hg_graph* g, hg_timing* out);

/* Phase timing for later hg_decompose calls (on != 0): an event is recorded
 * before every scan and every cascade launch, so hg_timing.scan_ms and
 * cascade_ms are filled in.  The events add a little device time to each
 * step (a few thousand steps on a 100M-vertex graph); off by default. */
int hg_set_phase_timing(hg_graph* g, int32_t on);

/* Graph facts for reporting. */
/* device_bytes: the handle's device memory as allocated (cache blocks, rounded up). */
int hg_graph_info(hg_graph* g, int64_t* n, int32_t* L, int64_t* directed_entries, int64_t* device_bytes);

int hg_graph_destroy(hg_graph* g);

const char* hg_last_error(void);

/* ------------------------------------------------------------------ batches
 * Many independent graphs decomposed in one persistent launch.  Every
 * (graph, lmd) pair is a work item in a device queue, ordered by estimated
 * cost; a fixed grid of workgroups pulls items and each workgroup runs its
 * item's whole k-chain.  The reference has no batched form (each of its
 * processes runs one method on one graph, main.cpp:229-257); per item the
 * algorithm and the results are FCSyncLeft::Execute's (FCSyncLeft.cpp:
 * 202-252), with the lmd-chains of all graphs run as concurrent tasks the
 * way ParGridPlus runs one graph's lmd-chains (FCSyncLeft.cpp:371-409).
 * Results are bit-exact against hg_decompose on each graph alone. */
typedef struct hg_batch hg_batch;

/* One graph of a batch: its layers as for hg_graph_create.  Limits per
 * graph: 1 <= L <= 64, 1 <= n, n * L < 2^32, fewer than 2^32 directed
 * entries. */
typedef struct hg_batch_graph {
    const hg_layer_csr* layers;
    int32_t L;
    int32_t offset_bits; /* 32 or 64 */
    int64_t n;
} hg_batch_graph;

/* Launch shape, chosen by hg_batch_tune or set by the caller. */
typedef struct hg_batch_params {
    int32_t block;         /* threads per workgroup: 256, 512 or 1024 */
    int32_t unroll;        /* entries per thread in flight per cascade round: 2 or 4 */
    int32_t blocks_per_cu; /* persistent workgroups per compute unit: 1, 2, 4 or 8 (capped at what
                              one CU holds resident for this block size) */
    int32_t order;         /* 0: items by estimated cost, largest first; 1: graph order */
    int32_t use_graph;     /* 1: replay a captured hipGraph; 0: plain launches */
    int32_t reserved;
} hg_batch_params;

/* Uploads all graphs (host memory) and allocates the state of every item.
 * stream: hipStream_t, or NULL for a stream the batch creates. */
int hg_batch_create(const hg_batch_graph* graphs, int32_t count, void* stream, hg_batch** out);

/* Sets (p != NULL) or resets to defaults (p == NULL) the launch shape. */
int hg_batch_set_params(hg_batch* b, const hg_batch_params* p);
int hg_batch_get_params(hg_batch* b, hg_batch_params* p);

/* Autotuner: times every candidate launch shape on this batch (one warm-up
 * and `reps` timed runs each, median), keeps the fastest and returns it
 * with its time.  Results are the same for every shape. */
int hg_batch_tune(hg_batch* b, int32_t reps, hg_batch_params* best, double* best_ms);

/* Decomposes every graph of the batch (all lmd) in one launch. */
int hg_batch_decompose(hg_batch* b);

/* Results of graph `graph` (0-based) after hg_batch_decompose: core index
 * [L][n] and cell lengths, as hg_core_index and hg_cells. */
int hg_batch_core_index(hg_batch* b, int32_t graph, int32_t* core_index);
int hg_batch_cells(hg_batch* b, int32_t graph, int64_t* cell_len, int32_t cap, int32_t* kmax);

/* Weighted densest subgraph (as hg_densest) of each listed graph, in one
 * launch; out[i] belongs to graphs[i]. */
int hg_batch_densest(hg_batch* b, const int32_t* graphs, int32_t count, double beta, hg_wds* out);

/* Device time of the last hg_batch_decompose / hg_batch_densest (ms). */
int hg_batch_last_ms(hg_batch* b, double* decompose_ms, double* densest_ms);

/* Batch facts: graphs, work items, directed entries, device bytes. */
int hg_batch_info(hg_batch* b, int32_t* count, int64_t* items, int64_t* directed_entries, int64_t* device_bytes);

int hg_batch_destroy(hg_batch* b);

/* ------------------------------------------------------------- live queue
 * A device work queue shared by batches: items of a batch submitted while
 * items of other batches are on the GPU are taken by the workgroups already
 * running (each takes the next item whenever it finishes one), so arrivals
 * join the running launch as its workgroups free up instead of waiting for
 * it to end.  A workgroup that finds the queue empty exits; every
 * submission also makes a launch of its own, so its items always run.
 * Per item the work and the results are those of hg_batch_decompose
 * (level 0 is pulled by the item's workgroup).  The reference has no
 * counterpart (one graph per process, main.cpp:229-257). */
typedef struct hg_live hg_live;

/* Creates a queue of `slots` item slots (<= 0: 65536) whose launches use
 * p's block, unroll and blocks_per_cu (NULL: defaults). */
int hg_live_create(const hg_batch_params* p, int32_t slots, hg_live** out);
/* Appends every item of b to the queue and returns; b's results are valid
 * after hg_batch_wait_live.  One submission at a time per batch; the batch
 * must not be decomposed or destroyed meanwhile (destroy waits). */
int hg_batch_submit_live(hg_batch* b, hg_live* q);
/* Blocks until b's submitted items are done; then hg_batch_core_index,
 * hg_batch_cells and hg_batch_densest work as after hg_batch_decompose,
 * and hg_batch_last_ms reports submit -> done wall time.  *taken_running
 * (may be NULL) = its items taken by launches that were already on the GPU
 * when b was submitted. */
int hg_batch_wait_live(hg_batch* b, int64_t* taken_running);
/* Items done, items taken by earlier launches, launches made. */
int hg_live_stats(hg_live* q, int64_t* items_done, int64_t* items_taken_running, int64_t* launches);
int hg_live_destroy(hg_live* q);

/* ------------------------------------------------------------------ serving
 * A request loop over one GPU.  Requests (a graph and "decompose" or
 * "densest") are copied in at submit and queue in arrival order; `lanes`
 * worker threads, each with its own HIP stream, take work when idle: a
 * request with more than `large_entries` directed entries runs alone on the
 * single-graph engine (hg_graph_*).  When at least `min_batch` small
 * requests wait at the head of the queue, a lane takes up to `max_batch` of
 * them and runs them as one batch (hg_batch_*, the throughput form); with
 * fewer waiting it takes one and runs it alone on the single-graph engine
 * (the latency form: the GPU is not saturated).  A request is admitted only
 * while the device bytes of all work in flight stay within `hbm_budget`,
 * counted as the library's device-memory cache hands the blocks out (each
 * rounded up to a power of two up to 512 MB, to 2 MB above); idle blocks the
 * cache keeps between requests (up to HG_POOL_MB) are outside the budget:
 * hg_server_create empties the cache, and a full device releases it before
 * an allocation fails.
 * Requests that arrive while others run join the next batch a lane forms;
 * the lanes' work overlaps on the device.  Every answer equals that of
 * hg_graph_* on the request's graph alone.
 * plain = 1 is the one-at-a-time baseline: one lane, every request alone on
 * the single-graph engine.  The reference has no serving layer (main.cpp
 * runs one method on one graph per process). */
typedef struct hg_server hg_server;

typedef struct hg_server_config {
    int64_t hbm_budget;     /* device bytes the loop may hold at once (registered graphs included) */
    int64_t large_entries;  /* requests above this many directed entries run alone */
    int32_t lanes;          /* concurrent workers (streams), 1..16 */
    int32_t max_batch;      /* requests per batch, >= 1 */
    int32_t plain;          /* 1: one request at a time on the single-graph engine */
    int32_t min_batch;      /* small requests waiting before a lane batches them (0: 8) */
    hg_batch_params params; /* launch shape of the batches (block == 0: defaults) */
    int32_t reuse_results;  /* 1: a registered graph's resident decomposition answers repeat
                               decompositions and densest searches on it; 0: every request
                               decomposes again */
    int32_t device_queue;   /* 1: small requests go, as soon as a lane takes them (up to
                               max_batch at once, no min_batch wait), to one device work queue
                               (hg_live) whose running workgroups take their items as they free
                               up; 0: host-formed batches, each its own launch */
} hg_server_config;

#define HG_REQ_DECOMPOSE 0
#define HG_REQ_DENSEST 1

typedef struct hg_request_info {
    int32_t status;      /* 0 done, 1 pending, < 0 the HG_ERR_* code it failed with */
    int32_t kind;        /* HG_REQ_* */
    int64_t n;
    int32_t L;
    int32_t batch_size;  /* requests in the batch it ran in (1: alone) */
    double queue_ms;     /* submit -> start of its batch */
    double run_ms;       /* start -> results on the host */
    double latency_ms;   /* submit -> results on the host */
    hg_wds wds;          /* HG_REQ_DENSEST: as hg_densest */
} hg_request_info;

typedef struct hg_server_summary {
    int64_t submitted, completed, failed, batches;
    double p50_ms, p99_ms, mean_ms, max_ms; /* latency of completed requests */
    double span_s;                          /* first submit -> last completion */
    int64_t peak_bytes;                     /* most device bytes admitted at once (registered graphs included) */
    int64_t cancelled;                      /* requests cancelled before they started */
    int64_t resident_graphs, resident_bytes; /* registered graphs and the device bytes they hold */
} hg_server_summary;

int hg_server_create(const hg_server_config* cfg, hg_server** out);
/* Queues a request (the graph is copied); *id names it. */
int hg_server_submit(hg_server* s, const hg_batch_graph* graph, int32_t kind, double beta, int64_t* id);
/* Registers a graph: it is uploaded and laid out once (on a stream of its
 * own) and stays resident in device memory, its bytes counted against
 * hbm_budget, until hg_server_unregister.  It is kept for the single-graph
 * engine, or, with device_queue and at most large_entries entries, as a
 * one-graph batch whose items join the device work queue.  Fails with
 * HG_ERR_NOMEM when the registered graphs would exceed the budget.  The
 * caller's arrays may be reused on return. */
int hg_server_register(hg_server* s, const hg_batch_graph* graph, int64_t* gid);
/* Queues a request on a registered graph: nothing is copied or uploaded.
 * Requests on one graph run one at a time on its resident handle; with
 * reuse_results the decomposition the handle holds answers repeat
 * decompositions, and densest searches only score its cells. */
int hg_server_submit_resident(hg_server* s, int64_t gid, int32_t kind, double beta, int64_t* id);
/* Frees a registered graph; HG_ERR_STATE while requests on it are queued or
 * running. */
int hg_server_unregister(hg_server* s, int64_t gid);
/* Cancels a request that has not started: it is done with status
 * HG_ERR_CANCELLED (hg_server_wait returns that code).  A started request
 * runs to the end (HG_ERR_STATE); cancelling a cancelled one is a no-op. */
int hg_server_cancel(hg_server* s, int64_t id);
/* Test hook: the after-th device allocation from now on (1: the next) fails
 * as if the device were full, in any handle of this process; 0 disarms it.
 * Lets tests show that a failed allocation fails only its own request. */
int hg_debug_fail_alloc(int64_t after);
/* Blocks until request id is done; fills info (may be NULL). */
int hg_server_wait(hg_server* s, int64_t id, hg_request_info* info);
/* Results of a done request: core index [L][n] and kmax [L] (either may be NULL). */
int hg_server_result(hg_server* s, int64_t id, int32_t* core_index, int32_t* kmax);
/* Forgets a done request and its results. */
int hg_server_release(hg_server* s, int64_t id);
/* Blocks until every submitted request is done (or cancelled). */
int hg_server_drain(hg_server* s);
int hg_server_stats(hg_server* s, hg_server_summary* out);
/* Finishes the queued work, stops the lanes and frees the server. */
int hg_server_destroy(hg_server* s);


#ifdef __cplusplus
}
#endif

#endif /* HIPGRID_H */
