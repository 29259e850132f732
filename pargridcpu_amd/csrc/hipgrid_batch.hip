// hipgrid_batch.hip -- C ABI of batched decomposition (hg_batch_* in include/hipgrid.h).
//
// Device memory of a batch of G graphs (graph g: n_g vertices, L_g layers,
// D_g directed entries; items = (g, lmd) pairs with lmd <= the largest
// number of layers in which one vertex of g has an edge):
//   off    uint32 [sum (n_g+1)]  per graph vertex-major adjacency offsets, local
//   adj    uint32 [sum D_g]      entry u*L_g + l (u: internal id in graph g)
//   adjz   uint8  [sum D_g]      z(u) of each entry's target
//   z      uint8  [sum n_g]      layers in which the vertex has an edge
//   deg0   uint32 [sum n_g*L_g]  initial layer degrees
//   perm   uint32 [sum n_g]      internal id -> caller's id
//   out    int32  [sum L_g*n_g]  core index per graph [lmd][v], caller's ids
// and per item, in pools: deg [n][L], cnt, idx, list, queue [n], liveness
// words [ceil(n/32)], and (allocated at the first densest call) the edge
// histogram [L][kcap].  The items (by id and in queue order) and the level-0
// pull tiles (graph, first vertex, count) sit in small arrays of their own.
// Internal ids sort each graph's vertices by total degree, descending.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "hipgrid.h"
#include "hipgrid_common.h"
// This unit launches only the set-up kernels of the single-graph engine.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wunused-function"
#include "hipgrid_batch_kernels.h"
#pragma clang diagnostic pop

using namespace hgb;
using hgc::dalloc;
using hgc::env_int;
using hgc::fail;

struct hg_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int cus = 0;
    int32_t count = 0;
    // per graph
    std::vector<uint32_t> n, L, kcap, max_nz;
    std::vector<uint64_t> D, vb, outb;
    std::vector<BGraph> graphs;
    // items
    std::vector<BItem> items;              // by item id: graph-major, lmd ascending
    std::vector<uint64_t> cost;
    std::vector<std::vector<int32_t>> item_of;  // [g][lmd-1] -> item id, -1 if trivial
    uint64_t totD = 0, hist_words = 0;
    int64_t device_bytes = 0;
    // device
    uint32_t* off = nullptr;
    uint32_t* adj = nullptr;
    uint8_t* adjz = nullptr;
    uint8_t* z = nullptr;
    uint32_t* deg0 = nullptr;
    uint32_t* perm = nullptr;
    int32_t* out = nullptr;
    BGraph* d_graphs = nullptr;
    BItem* d_items = nullptr;  // in queue order
    BItem* d_items_id = nullptr;  // by item id
    PullTile* d_tiles = nullptr;  // level-0 pull work of the pulled graphs
    // pull tiles: [0, nheavy) hubs, [nheavy, nmed) medium vertices, the rest light
    uint32_t ntiles = 0, nheavy = 0, nmed = 0, pull_nj = 0;
    uint32_t* d_head = nullptr;
    int32_t* d_kmax = nullptr;  // by item id
    unsigned long long* d_clock = nullptr;  // HG_ITEM_CLOCK=1 (diagnostic): item start/end wall clock
    uint32_t* degpool = nullptr;
    int32_t* cntpool = nullptr;
    int32_t* idxpool = nullptr;
    uint32_t* listpool = nullptr;
    uint32_t* qpool = nullptr;
    uint32_t* bitpool = nullptr;
    uint32_t* histpool = nullptr;
    BItem* d_sel = nullptr;  // items of a densest call
    size_t sel_cap = 0;
    // launch
    hg_batch_params params{};
    bool items_uploaded = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    // a submission to a live queue (hg_batch_submit_live) not yet waited for
    hg_live* live = nullptr;
    uint64_t live_first = 0, live_count = 0, live_instance = 0;
    std::chrono::steady_clock::time_point live_t0;
    int64_t live_taken_running = 0;  // of its items, those taken by launches made before it
    // results
    bool decomposed = false;
    std::vector<int32_t> h_kmax;  // by item id
    std::vector<BItem> h_queue;   // items in queue order (upload source)
    double decompose_ms = 0, densest_ms = 0;
};

namespace {

hg_batch_params default_params() {
    hg_batch_params p{};
    p.block = 512;
    p.unroll = 4;
    p.blocks_per_cu = 2;
    p.order = 0;
    p.use_graph = 1;
    return p;
}

bool params_valid(const hg_batch_params& p) {
    return (p.block == 256 || p.block == 512 || p.block == 1024) && (p.unroll == 2 || p.unroll == 4) &&
           (p.blocks_per_cu == 1 || p.blocks_per_cu == 2 || p.blocks_per_cu == 4 || p.blocks_per_cu == 8) &&
           (p.order == 0 || p.order == 1) && (p.use_graph == 0 || p.use_graph == 1);
}

void drop_graph(hg_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void release(hg_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);  // nothing queued may still use the memory
    drop_graph(b);
    void* ptrs[] = {b->off,     b->adj,     b->adjz,     b->z,        b->deg0,    b->perm,
                    b->out,     b->d_graphs, b->d_items, b->d_items_id, b->d_tiles, b->d_head,
                    b->d_kmax,  b->d_clock, b->degpool, b->cntpool,  b->idxpool,  b->listpool, b->qpool,
                    b->bitpool, b->histpool, b->d_sel};
    for (void* p : ptrs)
        if (p) hgc::dfree(p);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// Per-graph host preparation.
struct Prep {
    std::vector<std::vector<int64_t>> rp;  // int64 offsets per layer
    std::vector<uint32_t> perm, inv;
    std::vector<uint8_t> zc, zr;           // z in caller's / internal ids
    std::vector<int64_t> off64;            // merged offsets, internal ids
    std::vector<uint64_t> ent_z;           // [z]: entries of the vertices with z layers, then suffix sums
    uint32_t H = 0, Hp = 0, Hh = 0, max_nz = 0;
    uint64_t D = 0, maxdeg = 0;
};

int prepare(const hg_batch_graph& gr, int32_t gi, Prep& P) {
    const std::string tag = "graph " + std::to_string(gi) + ": ";
    if (!gr.layers) return fail(HG_ERR_ARG, tag + "layers is NULL");
    if (gr.L < 1 || gr.L > 64) return fail(HG_ERR_RANGE, tag + "L must be in [1, 64]");
    if (gr.n < 1) return fail(HG_ERR_ARG, tag + "n must be >= 1");
    if (uint64_t(gr.n) * uint64_t(gr.L) >= (1ull << 32)) return fail(HG_ERR_RANGE, tag + "n * L must be < 2^32");
    if (gr.offset_bits != 32 && gr.offset_bits != 64) return fail(HG_ERR_ARG, tag + "offset_bits must be 32 or 64");
    const int64_t n = gr.n;
    const int L = gr.L;
    P.rp.assign(size_t(L), {});
    for (int l = 0; l < L; ++l) {
        if (!gr.layers[l].row_ptr || !gr.layers[l].col) return fail(HG_ERR_ARG, tag + "layer pointer is NULL");
        std::vector<int64_t>& r = P.rp[size_t(l)];
        r.resize(size_t(n) + 1);
        if (gr.offset_bits == 64) {
            std::memcpy(r.data(), gr.layers[l].row_ptr, sizeof(int64_t) * (size_t(n) + 1));
        } else {
            const int32_t* s = static_cast<const int32_t*>(gr.layers[l].row_ptr);
            for (int64_t v = 0; v <= n; ++v) r[size_t(v)] = s[v];
        }
        if (r[0] != 0) return fail(HG_ERR_ARG, tag + "row_ptr[0] must be 0");
        for (int64_t v = 0; v < n; ++v) {
            if (r[size_t(v) + 1] < r[size_t(v)]) return fail(HG_ERR_ARG, tag + "row_ptr must be non-decreasing");
            P.maxdeg = std::max<uint64_t>(P.maxdeg, uint64_t(r[size_t(v) + 1] - r[size_t(v)]));
        }
    }
    std::vector<uint64_t> tot(static_cast<size_t>(n));
    P.zc.assign(size_t(n), 0);
    uint64_t maxtot = 0;
    for (int64_t v = 0; v < n; ++v) {
        uint64_t t = 0;
        uint32_t zz = 0;
        for (int l = 0; l < L; ++l) {
            const int64_t d = P.rp[size_t(l)][size_t(v) + 1] - P.rp[size_t(l)][size_t(v)];
            t += uint64_t(d);
            zz += d > 0;
        }
        tot[size_t(v)] = t;
        P.zc[size_t(v)] = uint8_t(zz);
        maxtot = std::max(maxtot, t);
        P.max_nz = std::max(P.max_nz, zz);
    }
    // entries of the vertices that take part in instance lmd (z >= lmd):
    // ent_z[lmd] after the suffix sums
    P.ent_z.assign(size_t(L) + 2, 0);
    for (int64_t v = 0; v < n; ++v) P.ent_z[P.zc[size_t(v)]] += tot[size_t(v)];
    for (int z = L - 1; z >= 0; --z) P.ent_z[size_t(z)] += P.ent_z[size_t(z) + 1];
    // internal id = rank by total degree, descending, stable in the caller's id
    P.perm.assign(size_t(n), 0);
    P.inv.assign(size_t(n), 0);
    {
        std::vector<uint64_t> bucket(size_t(maxtot) + 2, 0);
        for (int64_t v = 0; v < n; ++v) bucket[size_t(maxtot - tot[size_t(v)]) + 1]++;
        for (size_t q = 1; q < bucket.size(); ++q) bucket[q] += bucket[q - 1];
        for (int64_t v = 0; v < n; ++v) {
            const uint32_t r = uint32_t(bucket[size_t(maxtot - tot[size_t(v)])]++);
            P.perm[r] = uint32_t(v);
            P.inv[size_t(v)] = r;
        }
    }
    P.off64.assign(size_t(n) + 1, 0);
    P.zr.assign(size_t(n), 0);
    for (int64_t r = 0; r < n; ++r) {
        const uint32_t v = P.perm[size_t(r)];
        P.off64[size_t(r) + 1] = P.off64[size_t(r)] + int64_t(tot[v]);
        P.zr[size_t(r)] = P.zc[v];
        if (tot[v] > kWaveVertex) P.H = uint32_t(r) + 1;
        if (tot[v] > hg::kPullHeavy) P.Hp = uint32_t(r) + 1;
        if (tot[v] > kPullStageBytes) P.Hh = uint32_t(r) + 1;
    }
    P.D = uint64_t(P.off64[size_t(n)]);
    if (P.D >= (1ull << 32)) return fail(HG_ERR_RANGE, tag + "directed entries must be < 2^32");
    return HG_OK;
}

int create_impl(const hg_batch_graph* gs, int32_t count, void* stream, hg_batch* b) {
    if (!gs) return fail(HG_ERR_ARG, "graphs is NULL");
    if (count < 1) return fail(HG_ERR_ARG, "count must be >= 1");
    HIP_TRY(hipGetDevice(&b->device));
    HIP_TRY(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device));
    if (stream) {
        b->stream = static_cast<hipStream_t>(stream);
    } else {
        HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        b->own_stream = true;
    }
    for (auto& e : b->ev) HIP_TRY(hipEventCreate(&e));
    b->count = count;
    b->params = default_params();
    const size_t G = size_t(count);
    b->n.resize(G);
    b->L.resize(G);
    b->kcap.resize(G);
    b->max_nz.resize(G);
    b->D.resize(G);
    b->vb.resize(G);
    b->outb.resize(G);
    b->graphs.resize(G);
    b->item_of.resize(G);

    // Pass 1: validate and size everything.
    uint64_t V = 0, O = 0, E = 0, DB = 0, OUT = 0, RP = 0, LAY = 0;
    std::vector<uint64_t> ob(G), eb(G), db(G), rpb(G), layb(G);
    for (size_t g = 0; g < G; ++g) {
        const hg_batch_graph& gr = gs[g];
        if (gr.L < 1 || gr.L > 64) return fail(HG_ERR_RANGE, "graph " + std::to_string(g) + ": L must be in [1, 64]");
        if (gr.n < 1) return fail(HG_ERR_ARG, "graph " + std::to_string(g) + ": n must be >= 1");
        if (uint64_t(gr.n) * uint64_t(gr.L) >= (1ull << 32))
            return fail(HG_ERR_RANGE, "graph " + std::to_string(g) + ": n * L must be < 2^32");
        b->n[g] = uint32_t(gr.n);
        b->L[g] = uint32_t(gr.L);
        b->vb[g] = V;
        ob[g] = O;
        db[g] = DB;
        b->outb[g] = OUT;
        rpb[g] = RP;
        layb[g] = LAY;
        V += uint64_t(gr.n);
        O += uint64_t(gr.n) + 1;
        DB += uint64_t(gr.n) * uint64_t(gr.L);
        OUT += uint64_t(gr.n) * uint64_t(gr.L);
        RP += (uint64_t(gr.n) + 1) * uint64_t(gr.L);
        LAY += uint64_t(gr.L);
    }
    // Pass 2: per-graph host preparation (offsets, relabelling) and entry
    // counts; the device build needs every graph's offsets first.
    std::vector<Prep> prep(G);
    for (size_t g = 0; g < G; ++g) {
        int rc = prepare(gs[g], int32_t(g), prep[g]);
        if (rc) return rc;
        eb[g] = E;
        E += prep[g].D;
        b->D[g] = prep[g].D;
        b->max_nz[g] = prep[g].max_nz;
        b->kcap[g] = uint32_t(prep[g].maxdeg + 2);
    }
    b->totD = E;
    // Items: (g, lmd) for lmd <= max_nz; larger lmd have every index 0.
    uint64_t SDEG = 0, SN = 0, SW = 0, SH = 0;
    for (size_t g = 0; g < G; ++g) {
        BGraph& bg = b->graphs[g];
        bg.eb = eb[g];
        bg.vb = b->vb[g];
        bg.ob = ob[g];
        bg.db = db[g];
        bg.n = b->n[g];
        bg.L = b->L[g];
        bg.lmagic = bg.L == 1 ? 0xFFFFFFFFu : uint32_t((1ull << 32) / bg.L);
        bg.H = prep[g].H;
        bg.Hp = prep[g].Hp;
        bg.item0 = uint32_t(b->items.size());
        bg.nit = std::min(b->L[g], b->max_nz[g]);
        bg.pulled = bg.nit >= 1 && bg.nit <= hg::kPullMaxL;
        bg.hub16 = prep[g].maxdeg < (1u << 16);
        b->item_of[g].assign(b->L[g], -1);
        for (uint32_t lmd = 1; lmd <= b->L[g] && lmd <= b->max_nz[g]; ++lmd) {
            BItem it{};
            it.g = uint32_t(g);
            it.lmd = lmd;
            it.degb = SDEG;
            it.sb = SN;
            it.bitb = SW;
            it.outb = b->outb[g] + uint64_t(lmd - 1) * b->n[g];
            it.histb = SH;
            it.kcap = b->kcap[g];
            it.id = uint32_t(b->items.size());
            SDEG += uint64_t(b->n[g]) * b->L[g];
            SN += b->n[g];
            SW += (uint64_t(b->n[g]) + 31) / 32;
            SH += uint64_t(b->L[g]) * b->kcap[g];
            b->item_of[g][lmd - 1] = int32_t(b->items.size());
            b->items.push_back(it);
            // a chain visits the entries of its vertices (those with z >= lmd)
            // about once over all its levels, whatever lmd: measured chain
            // times of one graph differ by under 10 % across lmd
            b->cost.push_back(prep[g].ent_z[lmd] + uint64_t(b->n[g]));
        }
    }
    b->hist_words = SH;
    const size_t I = std::max<size_t>(b->items.size(), 1);
    // Level-0 pull work: a wavefront per vertex of the heavy head, a thread
    // per vertex for the rest, in tiles of one graph each.
    std::vector<PullTile> tiles;
    uint32_t max_pulled = 0;
    for (size_t g = 0; g < G; ++g) {
        const BGraph& bg = b->graphs[g];
        if (!bg.pulled) continue;
        max_pulled = std::max(max_pulled, bg.nit);
        for (uint32_t v = 0; v < prep[g].Hh; v += kPullHeavyTile)
            tiles.push_back(PullTile{uint32_t(g), v, std::min(kPullHeavyTile, prep[g].Hh - v), 1u});
    }
    b->nheavy = uint32_t(tiles.size());
    // whole vertices, at most kPullBlock rows and kPullStageBytes entries
    // (a vertex below Hh has at most kPullStageBytes entries)
    auto row_tiles = [&](size_t g, uint32_t v0, uint32_t v1) {
        const std::vector<int64_t>& o = prep[g].off64;
        const uint32_t vmax = kPullBlock / b->graphs[g].L;
        for (uint32_t v = v0; v < v1;) {
            uint32_t c = 1;
            while (v + c < v1 && c < vmax && uint64_t(o[v + c + 1] - o[v]) <= kPullStageBytes) ++c;
            tiles.push_back(PullTile{uint32_t(g), v, c, 0u});
            v += c;
        }
    };
    for (size_t g = 0; g < G; ++g)
        if (b->graphs[g].pulled) row_tiles(g, prep[g].Hh, b->graphs[g].Hp);
    b->nmed = uint32_t(tiles.size());
    for (size_t g = 0; g < G; ++g)
        if (b->graphs[g].pulled) row_tiles(g, b->graphs[g].Hp, b->graphs[g].n);
    b->ntiles = uint32_t(tiles.size());
    b->pull_nj = max_pulled <= 8 ? 8u : 16u;

    int64_t bytes = 0;
    auto acct = [&](uint64_t x) { bytes += int64_t(hgc::pool_block_bytes(x)); };  // as the cache hands blocks out
    HIP_TRY(dalloc(&b->off, O)); acct(4 * O);
    HIP_TRY(dalloc(&b->adj, E)); acct(4 * E);
    HIP_TRY(dalloc(&b->adjz, E + 16)); acct(E + 16);  // 16-byte reads past the last entry
    HIP_TRY(dalloc(&b->z, V)); acct(V);
    HIP_TRY(dalloc(&b->deg0, DB)); acct(4 * DB);
    HIP_TRY(dalloc(&b->perm, V)); acct(4 * V);
    HIP_TRY(dalloc(&b->out, OUT)); acct(4 * OUT);
    HIP_TRY(dalloc(&b->d_graphs, G));
    HIP_TRY(dalloc(&b->d_items, I));
    HIP_TRY(dalloc(&b->d_items_id, I));
    HIP_TRY(dalloc(&b->d_tiles, std::max<size_t>(tiles.size(), 1)));
    HIP_TRY(dalloc(&b->d_head, 1));
    HIP_TRY(dalloc(&b->d_kmax, I));
    if (env_int("HG_ITEM_CLOCK", 0) == 1) HIP_TRY(dalloc(&b->d_clock, 2 * std::max<size_t>(I, 1)));
    HIP_TRY(dalloc(&b->degpool, SDEG)); acct(4 * SDEG);
    HIP_TRY(dalloc(&b->cntpool, SN)); acct(4 * SN);
    HIP_TRY(dalloc(&b->idxpool, SN)); acct(4 * SN);
    HIP_TRY(dalloc(&b->listpool, SN)); acct(4 * SN);
    HIP_TRY(dalloc(&b->qpool, SN)); acct(4 * SN);
    HIP_TRY(dalloc(&b->bitpool, SW)); acct(4 * SW);
    b->device_bytes = bytes;
    if (env_int("HG_POISON", 0) == 1) {  // diagnostic: item state starts as all-ones words, not zeros
        HIP_TRY(hipMemsetAsync(b->degpool, 0xFF, 4 * std::max<uint64_t>(SDEG, 1), b->stream));
        HIP_TRY(hipMemsetAsync(b->cntpool, 0xFF, 4 * std::max<uint64_t>(SN, 1), b->stream));
        HIP_TRY(hipMemsetAsync(b->idxpool, 0xFF, 4 * std::max<uint64_t>(SN, 1), b->stream));
        HIP_TRY(hipMemsetAsync(b->listpool, 0xFF, 4 * std::max<uint64_t>(SN, 1), b->stream));
        HIP_TRY(hipMemsetAsync(b->qpool, 0xFF, 4 * std::max<uint64_t>(SN, 1), b->stream));
        HIP_TRY(hipMemsetAsync(b->bitpool, 0xFF, 4 * std::max<uint64_t>(SW, 1), b->stream));
    }

    // Upload and build: per-layer CSR into temporaries, then per graph the
    // single-graph engine's set-up kernels write deg0 and the merged,
    // relabelled adjacency.
    std::vector<uint32_t> h_off(O);
    std::vector<uint8_t> h_z(V);
    std::vector<uint32_t> h_perm(V), h_inv(V);
    std::vector<uint8_t> h_zc(V);
    std::vector<int64_t> h_off64(O);
    std::vector<int64_t> h_rp(RP);
    for (size_t g = 0; g < G; ++g) {
        const Prep& P = prep[g];
        const size_t n = b->n[g];
        for (size_t v = 0; v <= n; ++v) {
            h_off[ob[g] + v] = uint32_t(P.off64[v]);
            h_off64[ob[g] + v] = P.off64[v];
        }
        std::memcpy(&h_z[b->vb[g]], P.zr.data(), n);
        std::memcpy(&h_zc[b->vb[g]], P.zc.data(), n);
        std::memcpy(&h_perm[b->vb[g]], P.perm.data(), 4 * n);
        std::memcpy(&h_inv[b->vb[g]], P.inv.data(), 4 * n);
        for (uint32_t l = 0; l < b->L[g]; ++l)
            std::memcpy(&h_rp[rpb[g] + uint64_t(l) * (n + 1)], P.rp[l].data(), 8 * (n + 1));
    }
    int64_t* d_rp = nullptr;
    uint32_t* d_col = nullptr;
    int64_t** d_rp_ptr = nullptr;
    uint32_t** d_col_ptr = nullptr;
    int64_t* d_off64 = nullptr;
    uint32_t* d_inv = nullptr;
    uint8_t* d_zc = nullptr;
    unsigned long long* d_prep = nullptr;
    auto cleanup = [&]() {
        void* t[] = {d_rp, d_col, d_rp_ptr, d_col_ptr, d_off64, d_inv, d_zc, d_prep};
        for (void* p : t)
            if (p) hgc::dfree(p);
    };
    hipError_t e = dalloc(&d_rp, RP);
    if (e == hipSuccess) e = dalloc(&d_col, E);
    if (e == hipSuccess) e = dalloc(&d_rp_ptr, LAY);
    if (e == hipSuccess) e = dalloc(&d_col_ptr, LAY);
    if (e == hipSuccess) e = dalloc(&d_off64, O);
    if (e == hipSuccess) e = dalloc(&d_inv, V);
    if (e == hipSuccess) e = dalloc(&d_zc, V);
    if (e == hipSuccess) e = dalloc(&d_prep, size_t(hg::kPrepWords));
    if (e == hipSuccess) e = hipMemsetAsync(d_prep, 0, sizeof(unsigned long long) * hg::kPrepWords, b->stream);
    std::vector<int64_t*> rp_ptr(LAY);
    std::vector<uint32_t*> col_ptr(LAY);
    hgc::Staging up(b->stream);
    if (e == hipSuccess) {
        uint64_t cpos = 0;
        for (size_t g = 0; g < G; ++g) {
            const size_t n = b->n[g];
            for (uint32_t l = 0; l < b->L[g]; ++l) {
                rp_ptr[layb[g] + l] = d_rp + rpb[g] + uint64_t(l) * (n + 1);
                col_ptr[layb[g] + l] = d_col + cpos;
                const size_t m = size_t(prep[g].rp[l][n]);
                if (m && e == hipSuccess) e = up.h2d(d_col + cpos, gs[g].layers[l].col, 4 * m);
                cpos += m;
            }
        }
    }
    if (e == hipSuccess) e = up.h2d(d_rp, h_rp.data(), 8 * RP);
    if (e == hipSuccess) e = up.h2d(d_rp_ptr, rp_ptr.data(), sizeof(int64_t*) * LAY);
    if (e == hipSuccess) e = up.h2d(d_col_ptr, col_ptr.data(), sizeof(uint32_t*) * LAY);
    if (e == hipSuccess) e = up.h2d(d_off64, h_off64.data(), 8 * O);
    if (e == hipSuccess) e = up.h2d(d_inv, h_inv.data(), 4 * V);
    if (e == hipSuccess) e = up.h2d(d_zc, h_zc.data(), V);
    if (e == hipSuccess) e = up.h2d(b->off, h_off.data(), 4 * O);
    if (e == hipSuccess) e = up.h2d(b->z, h_z.data(), V);
    if (e == hipSuccess) e = up.h2d(b->perm, h_perm.data(), 4 * V);
    if (e == hipSuccess) e = hipMemsetAsync(b->out, 0, 4 * std::max<uint64_t>(OUT, 1), b->stream);
    if (e == hipSuccess) e = up.h2d(b->d_graphs, b->graphs.data(), sizeof(BGraph) * G);
    if (e == hipSuccess && !b->items.empty()) e = up.h2d(b->d_items_id, b->items.data(), sizeof(BItem) * b->items.size());
    if (e == hipSuccess && !tiles.empty()) e = up.h2d(b->d_tiles, tiles.data(), sizeof(PullTile) * tiles.size());
    if (e == hipSuccess) e = hipMemsetAsync(b->adjz + E, 0, 16, b->stream);
    for (size_t g = 0; g < G && e == hipSuccess; ++g) {
        const uint64_t nL = uint64_t(b->n[g]) * b->L[g];
        const uint32_t grid = uint32_t(std::min<uint64_t>((nL + hg::kBlock - 1) / hg::kBlock, uint64_t(b->cus) * 16));
        hipLaunchKernelGGL(hg::k_build_deg0, dim3(std::max(grid, 1u)), dim3(hg::kBlock), 0, b->stream,
                           d_rp_ptr + layb[g], b->perm + b->vb[g], b->n[g], b->L[g], b->deg0 + db[g]);
        const uint32_t grid2 = uint32_t(std::min<uint64_t>((uint64_t(b->n[g]) * kWave + hg::kBlock - 1) / hg::kBlock,
                                                           uint64_t(b->cus) * 16));
        hipLaunchKernelGGL(hg::k_build_adj, dim3(std::max(grid2, 1u)), dim3(hg::kBlock), 0, b->stream,
                           d_rp_ptr + layb[g], d_col_ptr + layb[g], b->deg0 + db[g], d_off64 + ob[g], b->perm + b->vb[g],
                           d_inv + b->vb[g], d_zc + b->vb[g], b->n[g], b->L[g], b->adj + eb[g], b->adjz + eb[g],
                           d_prep);
        e = hipGetLastError();
    }
    // also on failure: queued kernels read the temporaries freed below
    {
        const hipError_t es = hipStreamSynchronize(b->stream);
        if (e == hipSuccess) e = es;
    }
    unsigned long long h_prep[hg::kPrepWords] = {};
    if (e == hipSuccess) e = up.d2h(h_prep, d_prep, sizeof h_prep);
    if (e == hipSuccess && (h_prep[hg::kPrepBad] & hg::kBadCol)) {
        cleanup();
        return fail(HG_ERR_ARG, "a neighbour id is >= n in a batch graph");
    }
    if (e == hipSuccess && env_int("HG_VERIFY_UPLOAD", 0) == 1) {  // diagnostic: read every layer back
        std::vector<int64_t> brp(RP);
        e = up.d2h(brp.data(), d_rp, 8 * RP);
        bool same = e == hipSuccess && std::memcmp(brp.data(), h_rp.data(), 8 * RP) == 0;
        uint64_t cpos = 0;
        for (size_t g = 0; g < G && same && e == hipSuccess; ++g)
            for (uint32_t l = 0; l < b->L[g] && same && e == hipSuccess; ++l) {
                const size_t m = size_t(prep[g].rp[l][b->n[g]]);
                std::vector<uint32_t> bcol(m);
                if (m) e = up.d2h(bcol.data(), d_col + cpos, 4 * m);
                same = e == hipSuccess && (m == 0 || std::memcmp(bcol.data(), gs[g].layers[l].col, 4 * m) == 0);
                cpos += m;
            }
        if (e == hipSuccess && !same) {
            cleanup();
            return fail(HG_ERR_HIP, "upload verify: a graph's layers differ on the device");
        }
    }
    cleanup();
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? HG_ERR_NOMEM : HG_ERR_HIP, std::string("batch upload: ") + hipGetErrorString(e));
    return HG_OK;
}

int upload_items(hg_batch* b) {
    std::vector<BItem>& q = b->h_queue;  // lives in the handle (source of a queued copy)
    q = b->items;
    if (b->params.order == 0) {
        std::vector<size_t> ord(q.size());
        for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t c) { return b->cost[a] > b->cost[c]; });
        for (size_t i = 0; i < ord.size(); ++i) q[i] = b->items[ord[i]];
    }
    hgc::Staging up(b->stream);
    if (!q.empty()) HIP_TRY(up.h2d(b->d_items, q.data(), sizeof(BItem) * q.size()));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->items_uploaded = true;
    return HG_OK;
}

template <int B, int U>
void launch_peel_t(hg_batch* b, uint32_t grid) {
    hipLaunchKernelGGL((k_batch_peel<B, U>), dim3(grid), dim3(B), 0, b->stream, b->d_graphs, b->d_items,
                       uint32_t(b->items.size()), b->d_head, b->off, b->adj, b->adjz, b->z, b->deg0, b->perm,
                       b->degpool, b->cntpool, b->idxpool, b->listpool, b->qpool, b->bitpool, b->out, b->d_kmax,
                       b->d_clock);
}

template <int B, int U>
int resident_t() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&k_batch_peel<B, U>), B, 0) !=
        hipSuccess)
        return 1;
    return std::max(nb, 1);
}

// Workgroups of this launch shape that one CU holds at once (registers and
// LDS): more per CU would only queue behind the resident ones.
int resident_per_cu(int block, int unroll) {
    switch (block * 10 + unroll) {
        case 2562: return resident_t<256, 2>();
        case 2564: return resident_t<256, 4>();
        case 5122: return resident_t<512, 2>();
        case 5124: return resident_t<512, 4>();
        case 10242: return resident_t<1024, 2>();
        default: return resident_t<1024, 4>();
    }
}

void launch_peel(hg_batch* b) {
    const uint32_t cap = uint32_t(b->cus) * 16;
    const uint32_t nmed = b->nmed - b->nheavy, nlight = b->ntiles - b->nmed;
    const PullTile* med = b->d_tiles + b->nheavy;
    const PullTile* light = b->d_tiles + b->nmed;
    if (b->nheavy) {
        const uint32_t grid = std::min(b->nheavy, cap);
        if (b->pull_nj == 8)
            hipLaunchKernelGGL(k_batch_pull_heavy<8>, dim3(grid), dim3(kPullBlock), 0, b->stream, b->d_graphs,
                               b->d_items_id, b->d_tiles, b->nheavy, b->off, b->adjz, b->z, b->deg0, b->degpool,
                               b->cntpool);
        else
            hipLaunchKernelGGL(k_batch_pull_heavy<16>, dim3(grid), dim3(kPullBlock), 0, b->stream, b->d_graphs,
                               b->d_items_id, b->d_tiles, b->nheavy, b->off, b->adjz, b->z, b->deg0, b->degpool,
                               b->cntpool);
    }
    if (nmed) {
        const uint32_t grid = std::min(nmed, cap);
        if (b->pull_nj == 8)
            hipLaunchKernelGGL((k_batch_pull_light<8, 16>), dim3(grid), dim3(kPullBlock), 0, b->stream, b->d_graphs,
                               b->d_items_id, med, nmed, b->off, b->adjz, b->z, b->deg0, b->degpool, b->cntpool);
        else
            hipLaunchKernelGGL((k_batch_pull_light<16, 16>), dim3(grid), dim3(kPullBlock), 0, b->stream, b->d_graphs,
                               b->d_items_id, med, nmed, b->off, b->adjz, b->z, b->deg0, b->degpool, b->cntpool);
    }
    if (nlight) {
        const uint32_t grid = std::min(nlight, cap);
        if (b->pull_nj == 8)
            hipLaunchKernelGGL((k_batch_pull_light<8, 8>), dim3(grid), dim3(kPullBlock), 0, b->stream, b->d_graphs,
                               b->d_items_id, light, nlight, b->off, b->adjz, b->z, b->deg0, b->degpool, b->cntpool);
        else
            hipLaunchKernelGGL((k_batch_pull_light<16, 8>), dim3(grid), dim3(kPullBlock), 0, b->stream, b->d_graphs,
                               b->d_items_id, light, nlight, b->off, b->adjz, b->z, b->deg0, b->degpool, b->cntpool);
    }
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, b->stream, b->d_head);
    const int bpc = std::min(b->params.blocks_per_cu, resident_per_cu(b->params.block, b->params.unroll));
    const uint32_t want = uint32_t(b->cus) * uint32_t(std::max(bpc, 1));
    const uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(want, uint32_t(b->items.size())));
    switch (b->params.block * 10 + b->params.unroll) {
        case 2562: launch_peel_t<256, 2>(b, grid); break;
        case 2564: launch_peel_t<256, 4>(b, grid); break;
        case 5122: launch_peel_t<512, 2>(b, grid); break;
        case 5124: launch_peel_t<512, 4>(b, grid); break;
        case 10242: launch_peel_t<1024, 2>(b, grid); break;
        default: launch_peel_t<1024, 4>(b, grid); break;
    }
}

int decompose_impl(hg_batch* b) {
    HIP_TRY(hipSetDevice(b->device));
    b->decomposed = false;
    if (!b->items_uploaded) {
        int rc = upload_items(b);
        if (rc) return rc;
    }
    if (b->items.empty()) {  // no graph has an edge: every index is 0
        b->h_kmax.clear();
        b->decompose_ms = 0;
        b->decomposed = true;
        return HG_OK;
    }
    if (b->params.use_graph && !b->exec) {
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        launch_peel(b);
        hipGraph_t gr = nullptr;
        const hipError_t ec = hipStreamEndCapture(b->stream, &gr);
        if (ec != hipSuccess) return fail(HG_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(ec));
        b->graph = gr;
        HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    if (b->params.use_graph) {
        HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    } else {
        launch_peel(b);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    HIP_TRY(hipEventSynchronize(b->ev[1]));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    b->decompose_ms = ms;
    b->h_kmax.assign(b->items.size(), 0);
    hgc::Staging io(b->stream);
    HIP_TRY(io.d2h(b->h_kmax.data(), b->d_kmax, sizeof(int32_t) * b->items.size()));
    if (b->d_clock) {  // diagnostic: item durations and finish times (wall clock of the device)
        std::vector<unsigned long long> c(2 * b->items.size());
        HIP_TRY(io.d2h(c.data(), b->d_clock, sizeof(unsigned long long) * c.size()));
        int khz = 100000;
        (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device);
        unsigned long long t0 = ~0ull, t1 = 0;
        for (size_t i = 0; i < b->items.size(); ++i) {
            t0 = std::min(t0, c[2 * i]);
            t1 = std::max(t1, c[2 * i + 1]);
        }
        std::vector<std::pair<double, size_t>> dur;
        for (size_t i = 0; i < b->items.size(); ++i) dur.emplace_back(double(c[2 * i + 1] - c[2 * i]) / khz, i);
        std::sort(dur.rbegin(), dur.rend());
        double sum = 0;
        for (auto& d : dur) sum += d.first;
        std::fprintf(stderr, "[hg batch] items %zu span %.3f ms sum of item times %.3f ms (%.1f per workgroup of %d)\n",
                     dur.size(), double(t1 - t0) / khz, sum, sum / std::max(1, b->cus * std::max(1, b->params.blocks_per_cu)),
                     b->cus * b->params.blocks_per_cu);
        for (size_t j = 0; j < std::min<size_t>(dur.size(), 20); ++j) {
            const size_t i = dur[j].second;
            std::fprintf(stderr, "[hg batch]   item %zu graph %u lmd %u n %u L %u: %.3f ms, starts %.3f ends %.3f\n", i,
                         b->items[i].g, b->items[i].lmd, b->graphs[b->items[i].g].n, b->graphs[b->items[i].g].L,
                         dur[j].first, double(c[2 * i] - t0) / khz, double(c[2 * i + 1] - t0) / khz);
        }
    }
    b->decomposed = true;
    return HG_OK;
}

int check_graph(hg_batch* b, int32_t graph) {
    if (!b) return fail(HG_ERR_ARG, "batch is NULL");
    if (graph < 0 || graph >= b->count) return fail(HG_ERR_ARG, "graph index out of range");
    if (!b->decomposed) return fail(HG_ERR_STATE, "hg_batch_decompose has not run");
    return HG_OK;
}

int32_t kmax_of(const hg_batch* b, int32_t g, uint32_t lmd) {
    const int32_t it = b->item_of[size_t(g)][lmd - 1];
    return it < 0 ? 0 : b->h_kmax[size_t(it)];
}

// FCPathLevelLeft::constructCore's wds block (Core/FCPathLevelLeft.cpp:
// 57-99) for one cell: num_edge[l] are uint32 sums, length = |S(k, lmd)|.
void score_cell(const uint32_t* num_edge, uint32_t L, uint32_t length, const double* pw, double* density,
                int32_t* sel) {
    double dl[64];
    for (uint32_t l = 0; l < L; ++l) dl[l] = double(float(num_edge[l]) * 1.0f / 2);
    std::sort(dl, dl + L, [](double a, double c) { return a > c; });
    double best = 0.0;
    int32_t bsel = 0;
    for (uint32_t l = 0; l < L; ++l) {
        const double mad = (dl[l] * 1.0f) / double(length) * 1.0;
        const double dens = mad * pw[l];
        if (dens >= best) {
            best = dens;
            bsel = int32_t(l + 1);
        }
    }
    *density = best;
    *sel = bsel;
}

int densest_impl(hg_batch* b, const int32_t* graphs, int32_t count, double beta, hg_wds* out) {
    HIP_TRY(hipSetDevice(b->device));
    if (!b->histpool) {
        hipError_t e = dalloc(&b->histpool, std::max<uint64_t>(b->hist_words, 1));
        if (e != hipSuccess) return fail(HG_ERR_NOMEM, "batch densest histogram allocation failed");
        b->device_bytes += int64_t(hgc::pool_block_bytes(4 * b->hist_words));
    }
    std::vector<uint32_t> sel;
    for (int32_t i = 0; i < count; ++i) {
        const int32_t g = graphs[i];
        for (uint32_t lmd = 1; lmd <= b->L[size_t(g)]; ++lmd) {
            const int32_t it = b->item_of[size_t(g)][lmd - 1];
            if (it >= 0 && b->h_kmax[size_t(it)] >= 1) sel.push_back(uint32_t(it));
        }
    }
    std::sort(sel.begin(), sel.end());
    sel.erase(std::unique(sel.begin(), sel.end()), sel.end());
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    hgc::Staging io(b->stream);
    std::vector<BItem> sitems(sel.size());
    if (!sel.empty()) {
        if (sel.size() > b->sel_cap) {
            if (b->d_sel) {
                (void)hipStreamSynchronize(b->stream);
                hgc::dfree(b->d_sel);
            }
            b->d_sel = nullptr;
            b->sel_cap = 0;
            HIP_TRY(dalloc(&b->d_sel, sel.size()));
            b->sel_cap = sel.size();
        }
        for (size_t i = 0; i < sel.size(); ++i) {
            sitems[i] = b->items[sel[i]];
            const BItem& it = sitems[i];
            HIP_TRY(hipMemsetAsync(b->histpool + it.histb, 0, 4 * uint64_t(b->L[it.g]) * it.kcap, b->stream));
        }
        HIP_TRY(io.h2d(b->d_sel, sitems.data(), sizeof(BItem) * sel.size()));
        const uint32_t grid = uint32_t(std::min<size_t>(sel.size(), size_t(b->cus) * 4));
        hipLaunchKernelGGL((k_batch_edge_hist<256>), dim3(grid), dim3(256), 0, b->stream, b->d_graphs, b->d_sel,
                           uint32_t(sel.size()), b->off, b->adj, b->idxpool, b->histpool);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    HIP_TRY(hipEventSynchronize(b->ev[1]));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    b->densest_ms = ms;
    // Score on the host: cell lengths from the core-index rows, per-layer
    // edge counts from suffix sums of the histogram.
    for (int32_t i = 0; i < count; ++i) {
        const int32_t g = graphs[i];
        const uint32_t n = b->n[size_t(g)], L = b->L[size_t(g)];
        std::vector<double> pw(L);
        for (uint32_t l = 0; l < L; ++l) pw[l] = std::pow(double(l + 1), beta);
        std::vector<int32_t> ci(size_t(n) * L);
        HIP_TRY(io.d2h(ci.data(), b->out + b->outb[size_t(g)], 4 * ci.size()));
        hg_wds r{};
        float maxd = 0.0f;
        for (uint32_t lmd = 1; lmd <= L; ++lmd) {
            const int32_t km = kmax_of(b, g, lmd);
            if (km < 1) continue;
            const BItem& it = b->items[size_t(b->item_of[size_t(g)][lmd - 1])];
            std::vector<uint32_t> h(size_t(L) * it.kcap);
            HIP_TRY(io.d2h(h.data(), b->histpool + it.histb, 4 * h.size()));
            std::vector<uint64_t> cnt(size_t(km) + 2, 0);
            for (uint32_t v = 0; v < n; ++v) {
                const int32_t x = ci[size_t(lmd - 1) * n + v];
                if (x < 0 || x > km)
                    return fail(HG_ERR_DEVICE, "core index " + std::to_string(x) + " outside [0, " + std::to_string(km) +
                                                   "] in batch graph " + std::to_string(g) + " (n " + std::to_string(n) +
                                                   ", L " + std::to_string(L) + ", lmd " + std::to_string(lmd) + ")");
                cnt[size_t(x)]++;
            }
            std::vector<uint32_t> acc(L, 0);  // suffix sums over m >= k, uint32 as the reference
            uint64_t len = 0;
            std::vector<double> dens(size_t(km) + 1);
            std::vector<int32_t> dsel(size_t(km) + 1);
            std::vector<uint32_t> lens(size_t(km) + 1);
            for (int32_t k = km; k >= 1; --k) {
                for (uint32_t l = 0; l < L; ++l) acc[l] += h[size_t(l) * it.kcap + size_t(k)];
                len += cnt[size_t(k)];
                lens[size_t(k)] = uint32_t(len);
                score_cell(acc.data(), L, uint32_t(len), pw.data(), &dens[size_t(k)], &dsel[size_t(k)]);
            }
            // argmax as FCTree::WeightDenestSubgraph with main.cpp's float
            // running max: lmd-major, k ascending, update on >=
            for (int32_t k = 1; k <= km; ++k) {
                if (dens[size_t(k)] >= double(maxd)) {
                    maxd = float(dens[size_t(k)]);
                    r.res_k = uint32_t(k);
                    r.res_lmd = lmd;
                    r.res_len = lens[size_t(k)];
                    r.num_select_layer = dsel[size_t(k)];
                }
            }
        }
        r.max_density = maxd;
        out[i] = r;
    }
    return HG_OK;
}

}  // namespace

// A device work queue shared by batches (see k_live_peel): a ring of slots
// in device memory, head/tail words, done flags in host memory that the
// workgroups raise, and the streams its launches and copies go on.
struct hg_live {
    int device = 0;
    int cus = 0;
    hg_batch_params params{};
    uint32_t cap = 0;  // ring slots
    LiveCtl* ctl = nullptr;
    LiveSlot* ring = nullptr;
    unsigned long long* flags_h = nullptr;  // [cap][2]: done position + 1, the launch that took it
    unsigned long long* flags_d = nullptr;
    hipStream_t copy = nullptr;
    std::vector<hipStream_t> ks;  // launch streams, used in turn
    size_t next_ks = 0;
    hipEvent_t published = nullptr;
    std::mutex mu;
    uint64_t tail = 0;      // positions handed out
    uint64_t instance = 0;  // launches made
    std::vector<LiveSlot> staging;
    int64_t items_done = 0, items_taken_running = 0;
};

namespace {

void live_release(hg_live* q) {
    if (!q) return;
    (void)hipSetDevice(q->device);
    for (hipStream_t st : q->ks)
        if (st) (void)hipStreamSynchronize(st);
    if (q->copy) (void)hipStreamSynchronize(q->copy);
    if (q->ctl) hgc::dfree(q->ctl);
    if (q->ring) hgc::dfree(q->ring);
    if (q->flags_h) (void)hipHostFree(q->flags_h);
    for (hipStream_t st : q->ks)
        if (st) (void)hipStreamDestroy(st);
    if (q->copy) (void)hipStreamDestroy(q->copy);
    if (q->published) (void)hipEventDestroy(q->published);
    delete q;
}

template <int B, int U>
void launch_live_t(hg_live* q, hipStream_t st, uint32_t grid, unsigned long long inst) {
    hipLaunchKernelGGL((k_live_peel<B, U>), dim3(grid), dim3(B), 0, st, q->ctl, q->ring, q->cap, q->flags_d, inst);
}

template <int B, int U>
int live_resident_t() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&k_live_peel<B, U>), B, 0) !=
        hipSuccess)
        return 1;
    return std::max(nb, 1);
}

int live_resident_per_cu(int block, int unroll) {
    switch (block * 10 + unroll) {
        case 2562: return live_resident_t<256, 2>();
        case 2564: return live_resident_t<256, 4>();
        case 5122: return live_resident_t<512, 2>();
        case 5124: return live_resident_t<512, 4>();
        case 10242: return live_resident_t<1024, 2>();
        default: return live_resident_t<1024, 4>();
    }
}

volatile unsigned long long* live_flag(hg_live* q, uint64_t pos) {
    return reinterpret_cast<volatile unsigned long long*>(q->flags_h + 2 * (pos % q->cap));
}

int submit_live_impl(hg_batch* b, hg_live* q) {
    HIP_TRY(hipSetDevice(b->device));
    if (b->device != q->device) return fail(HG_ERR_ARG, "batch and live queue are on different devices");
    if (b->live) return fail(HG_ERR_STATE, "the batch has a submission not yet waited for");
    b->decomposed = false;
    if (!b->items_uploaded) {
        int rc = upload_items(b);
        if (rc) return rc;
    }
    const uint64_t n = b->h_queue.size();
    b->live_t0 = std::chrono::steady_clock::now();
    b->live = q;
    b->live_count = n;
    b->live_taken_running = 0;
    if (n == 0) {
        b->live_first = 0;
        return HG_OK;
    }
    if (n > q->cap) return fail(HG_ERR_RANGE, "batch has more items than the live queue has slots");
    // the batch's memory is written on its own stream: settled before any workgroup reads it
    HIP_TRY(hipStreamSynchronize(b->stream));
    std::lock_guard<std::mutex> lk(q->mu);
    const uint64_t first = q->tail;
    // slots are reused once the item that held them cap positions ago is done
    for (uint64_t p = first; p < first + n; ++p)
        if (p >= q->cap) {
            const uint64_t old = p - q->cap;
            const auto t0 = std::chrono::steady_clock::now();
            while (*live_flag(q, old) != old + 1) {
                if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(300))
                    return fail(HG_ERR_DEVICE, "live queue: a slot was not released within 300 s");
                std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
        }
    const unsigned long long inst = ++q->instance;
    q->staging.resize(n);
    const BPools P{b->off,     b->adj,      b->adjz,  b->z,       b->deg0, b->perm,  b->degpool, b->cntpool,
                   b->idxpool, b->listpool, b->qpool, b->bitpool, b->out,  b->d_kmax, nullptr};
    for (uint64_t i = 0; i < n; ++i) {
        LiveSlot& sl = q->staging[i];
        std::memset(&sl, 0, sizeof sl);
        sl.seq = first + i;
        sl.instance = inst;
        sl.P = P;
        sl.item = b->h_queue[i];
        sl.gr = b->graphs[sl.item.g];
        volatile unsigned long long* f = live_flag(q, first + i);
        f[0] = 0;
        f[1] = 0;
    }
    hgc::Staging io(q->copy);
    for (uint64_t i = 0; i < n;) {  // the ring may wrap
        const uint64_t slot = (first + i) % q->cap;
        const uint64_t m = std::min<uint64_t>(n - i, q->cap - slot);
        HIP_TRY(io.h2d(q->ring + slot, q->staging.data() + i, sizeof(LiveSlot) * m));
        i += m;
    }
    const unsigned long long new_tail = first + n;
    HIP_TRY(io.h2d(&q->ctl->tail, &new_tail, sizeof new_tail));
    HIP_TRY(hipEventRecord(q->published, q->copy));
    hipStream_t st = q->ks[q->next_ks++ % q->ks.size()];
    HIP_TRY(hipStreamWaitEvent(st, q->published, 0));
    const int bpc = std::min(q->params.blocks_per_cu, live_resident_per_cu(q->params.block, q->params.unroll));
    const uint32_t grid = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(n, uint64_t(q->cus) * uint64_t(std::max(bpc, 1)))));
    switch (q->params.block * 10 + q->params.unroll) {
        case 2562: launch_live_t<256, 2>(q, st, grid, inst); break;
        case 2564: launch_live_t<256, 4>(q, st, grid, inst); break;
        case 5122: launch_live_t<512, 2>(q, st, grid, inst); break;
        case 5124: launch_live_t<512, 4>(q, st, grid, inst); break;
        case 10242: launch_live_t<1024, 2>(q, st, grid, inst); break;
        default: launch_live_t<1024, 4>(q, st, grid, inst); break;
    }
    HIP_TRY(hipGetLastError());
    q->tail = new_tail;
    b->live_first = first;
    b->live_instance = inst;
    return HG_OK;
}

int wait_live_impl(hg_batch* b) {
    hg_live* q = b->live;
    if (!q) return fail(HG_ERR_STATE, "the batch has no live submission");
    HIP_TRY(hipSetDevice(b->device));
    const auto t0 = std::chrono::steady_clock::now();
    int64_t running = 0;
    for (uint64_t p = b->live_first; p < b->live_first + b->live_count; ++p) {
        volatile unsigned long long* f = live_flag(q, p);
        while (f[0] != p + 1) {
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(300)) {
                b->live = nullptr;
                return fail(HG_ERR_DEVICE, "live queue: the batch's items did not finish within 300 s");
            }
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
        running += f[1] < b->live_instance;  // taken by a launch already on the GPU at submission
    }
    b->live = nullptr;
    int32_t err = 0;
    HIP_TRY(hipMemcpy(&err, &q->ctl->error, sizeof err, hipMemcpyDeviceToHost));
    if (err) return fail(HG_ERR_DEVICE, "live queue: a slot did not hold the item it was claimed for");
    {
        std::lock_guard<std::mutex> lk(q->mu);
        q->items_done += int64_t(b->live_count);
        q->items_taken_running += running;
    }
    b->live_taken_running = running;
    b->h_kmax.assign(b->items.size(), 0);
    if (!b->items.empty()) {
        hgc::Staging io(b->stream);
        HIP_TRY(io.d2h(b->h_kmax.data(), b->d_kmax, sizeof(int32_t) * b->items.size()));
    }
    b->decompose_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - b->live_t0).count();
    b->decomposed = true;
    return HG_OK;
}

}  // namespace

extern "C" {

int hg_live_create(const hg_batch_params* p, int32_t slots, hg_live** out) {
    if (!out) return fail(HG_ERR_ARG, "out is NULL");
    *out = nullptr;
    hg_batch_params np = p ? *p : default_params();
    np.use_graph = 0;
    np.order = 0;
    if (!params_valid(np)) return fail(HG_ERR_ARG, "invalid launch parameters");
    if (slots < 1) slots = 65536;
    hg_live* q = new (std::nothrow) hg_live();
    if (!q) return fail(HG_ERR_NOMEM, "host allocation failed");
    q->params = np;
    q->cap = uint32_t(slots);
    auto bail = [&](int rc) {
        live_release(q);
        return rc;
    };
    if (hipGetDevice(&q->device) != hipSuccess ||
        hipDeviceGetAttribute(&q->cus, hipDeviceAttributeMultiprocessorCount, q->device) != hipSuccess)
        return bail(fail(HG_ERR_HIP, "no device"));
    if (dalloc(&q->ctl, 1) != hipSuccess || dalloc(&q->ring, size_t(q->cap)) != hipSuccess)
        return bail(fail(HG_ERR_NOMEM, "live queue: device allocation failed"));
    if (hipHostMalloc(reinterpret_cast<void**>(&q->flags_h), sizeof(unsigned long long) * 2 * q->cap,
                      hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess)
        return bail(fail(HG_ERR_NOMEM, "live queue: host allocation failed"));
    std::memset(q->flags_h, 0, sizeof(unsigned long long) * 2 * q->cap);
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&q->flags_d), q->flags_h, 0) != hipSuccess)
        return bail(fail(HG_ERR_HIP, "live queue: no device address for the done flags"));
    if (hipStreamCreateWithFlags(&q->copy, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&q->published, hipEventDisableTiming) != hipSuccess)
        return bail(fail(HG_ERR_HIP, "live queue: stream creation failed"));
    q->ks.assign(4, nullptr);
    for (auto& st : q->ks)
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
            return bail(fail(HG_ERR_HIP, "live queue: stream creation failed"));
    if (hipMemsetAsync(q->ctl, 0, sizeof(LiveCtl), q->copy) != hipSuccess || hipStreamSynchronize(q->copy) != hipSuccess)
        return bail(fail(HG_ERR_HIP, "live queue: initialisation failed"));
    *out = q;
    return HG_OK;
}

int hg_batch_submit_live(hg_batch* b, hg_live* q) {
    if (!b || !q) return fail(HG_ERR_ARG, "NULL argument");
    try {
        return submit_live_impl(b, q);
    } catch (const std::bad_alloc&) {
        b->live = nullptr;
        return fail(HG_ERR_NOMEM, "host allocation failed");
    }
}

int hg_batch_wait_live(hg_batch* b, int64_t* taken_running) {
    if (!b) return fail(HG_ERR_ARG, "batch is NULL");
    int rc;
    try {
        rc = wait_live_impl(b);
    } catch (const std::bad_alloc&) {
        rc = fail(HG_ERR_NOMEM, "host allocation failed");
    }
    if (taken_running) *taken_running = b->live_taken_running;
    return rc;
}

int hg_live_stats(hg_live* q, int64_t* items_done, int64_t* items_taken_running, int64_t* launches) {
    if (!q) return fail(HG_ERR_ARG, "live queue is NULL");
    std::lock_guard<std::mutex> lk(q->mu);
    if (items_done) *items_done = q->items_done;
    if (items_taken_running) *items_taken_running = q->items_taken_running;
    if (launches) *launches = int64_t(q->instance);
    return HG_OK;
}

int hg_live_destroy(hg_live* q) {
    if (!q) return fail(HG_ERR_ARG, "live queue is NULL");
    live_release(q);
    return HG_OK;
}

int hg_batch_create(const hg_batch_graph* graphs, int32_t count, void* stream, hg_batch** out) {
    if (!out) return fail(HG_ERR_ARG, "out is NULL");
    *out = nullptr;
    hg_batch* b = new (std::nothrow) hg_batch();
    if (!b) return fail(HG_ERR_NOMEM, "host allocation failed");
    int rc = HG_OK;
    try {
        rc = create_impl(graphs, count, stream, b);
    } catch (const std::bad_alloc&) {
        rc = fail(HG_ERR_NOMEM, "host allocation failed");
    }
    if (rc != HG_OK) {
        release(b);
        return rc;
    }
    *out = b;
    return HG_OK;
}

int hg_batch_set_params(hg_batch* b, const hg_batch_params* p) {
    if (!b) return fail(HG_ERR_ARG, "batch is NULL");
    const hg_batch_params np = p ? *p : default_params();
    if (!params_valid(np)) return fail(HG_ERR_ARG, "invalid batch launch parameters");
    HIP_TRY(hipSetDevice(b->device));
    const bool reorder = np.order != b->params.order;
    b->params = np;
    drop_graph(b);
    if (reorder) b->items_uploaded = false;
    return HG_OK;
}

int hg_batch_get_params(hg_batch* b, hg_batch_params* p) {
    if (!b || !p) return fail(HG_ERR_ARG, "NULL argument");
    *p = b->params;
    return HG_OK;
}

int hg_batch_tune(hg_batch* b, int32_t reps, hg_batch_params* best, double* best_ms) {
    if (!b) return fail(HG_ERR_ARG, "batch is NULL");
    if (reps < 1) reps = 1;
    const int blocks[] = {256, 512, 1024};
    const int unrolls[] = {2, 4};
    const int bpcs[] = {1, 2, 4, 8};
    hg_batch_params top = b->params;
    double top_ms = -1;
    try {
        for (int blk : blocks)
            for (int u : unrolls)
                for (int bpc : bpcs) {
                    if (blk * bpc > 2048) continue;  // beyond one CU's resident threads
                    if (bpc > resident_per_cu(blk, u)) continue;  // would not all be resident
                    hg_batch_params p = default_params();
                    p.block = blk;
                    p.unroll = u;
                    p.blocks_per_cu = bpc;
                    int rc = hg_batch_set_params(b, &p);
                    if (rc) return rc;
                    rc = decompose_impl(b);  // warm-up (and capture)
                    if (rc) return rc;
                    std::vector<double> t;
                    for (int r = 0; r < reps; ++r) {
                        rc = decompose_impl(b);
                        if (rc) return rc;
                        t.push_back(b->decompose_ms);
                    }
                    std::sort(t.begin(), t.end());
                    const double med = t[t.size() / 2];
                    if (top_ms < 0 || med < top_ms) {
                        top_ms = med;
                        top = p;
                    }
                }
        // item order, at the best shape
        hg_batch_params p = top;
        p.order = 1 - top.order;
        int rc = hg_batch_set_params(b, &p);
        if (rc) return rc;
        rc = decompose_impl(b);
        if (rc) return rc;
        std::vector<double> t;
        for (int r = 0; r < reps; ++r) {
            rc = decompose_impl(b);
            if (rc) return rc;
            t.push_back(b->decompose_ms);
        }
        std::sort(t.begin(), t.end());
        if (t[t.size() / 2] < top_ms) {
            top_ms = t[t.size() / 2];
            top = p;
        }
        rc = hg_batch_set_params(b, &top);
        if (rc) return rc;
        rc = decompose_impl(b);
        if (rc) return rc;
    } catch (const std::bad_alloc&) {
        return fail(HG_ERR_NOMEM, "host allocation failed");
    }
    if (best) *best = top;
    if (best_ms) *best_ms = top_ms;
    return HG_OK;
}

int hg_batch_decompose(hg_batch* b) {
    if (!b) return fail(HG_ERR_ARG, "batch is NULL");
    try {
        return decompose_impl(b);
    } catch (const std::bad_alloc&) {
        return fail(HG_ERR_NOMEM, "host allocation failed");
    }
}

int hg_batch_core_index(hg_batch* b, int32_t graph, int32_t* core_index) {
    int rc = check_graph(b, graph);
    if (rc) return rc;
    if (!core_index) return fail(HG_ERR_ARG, "core_index is NULL");
    HIP_TRY(hipSetDevice(b->device));
    hgc::Staging io(b->stream);
    HIP_TRY(io.d2h(core_index, b->out + b->outb[size_t(graph)],
                   sizeof(int32_t) * uint64_t(b->n[size_t(graph)]) * b->L[size_t(graph)]));
    return HG_OK;
}

int hg_batch_cells(hg_batch* b, int32_t graph, int64_t* cell_len, int32_t cap, int32_t* kmax) {
    int rc = check_graph(b, graph);
    if (rc) return rc;
    const uint32_t n = b->n[size_t(graph)], L = b->L[size_t(graph)];
    if (kmax)
        for (uint32_t lmd = 1; lmd <= L; ++lmd) kmax[lmd - 1] = kmax_of(b, graph, lmd);
    if (!cell_len) return HG_OK;
    if (cap < 1) return fail(HG_ERR_ARG, "cap must be >= 1");
    try {
        std::vector<int32_t> ci(size_t(n) * L);
        HIP_TRY(hipSetDevice(b->device));
        hgc::Staging io(b->stream);
        HIP_TRY(io.d2h(ci.data(), b->out + b->outb[size_t(graph)], 4 * ci.size()));
        for (uint32_t i = 0; i < L; ++i) {
            std::vector<int64_t> h(size_t(cap) + 1, 0);
            for (uint32_t v = 0; v < n; ++v) {
                const int32_t x = ci[size_t(i) * n + v];
                if (x < 0) return fail(HG_ERR_DEVICE, "negative core index in a batch result");
                h[size_t(std::min<int64_t>(x, cap))]++;
            }
            int64_t acc = 0;
            for (int32_t k = cap; k >= 0; --k) {
                acc += h[size_t(k)];
                if (k < cap) cell_len[size_t(i) * size_t(cap) + size_t(k)] = acc;
            }
        }
    } catch (const std::bad_alloc&) {
        return fail(HG_ERR_NOMEM, "host allocation failed");
    }
    return HG_OK;
}

int hg_batch_densest(hg_batch* b, const int32_t* graphs, int32_t count, double beta, hg_wds* out) {
    if (!b || !graphs || !out) return fail(HG_ERR_ARG, "NULL argument");
    if (count < 0) return fail(HG_ERR_ARG, "count must be >= 0");
    for (int32_t i = 0; i < count; ++i) {
        int rc = check_graph(b, graphs[i]);
        if (rc) return rc;
    }
    try {
        return densest_impl(b, graphs, count, beta, out);
    } catch (const std::bad_alloc&) {
        return fail(HG_ERR_NOMEM, "host allocation failed");
    }
}

int hg_batch_last_ms(hg_batch* b, double* decompose_ms, double* densest_ms) {
    if (!b) return fail(HG_ERR_ARG, "batch is NULL");
    if (decompose_ms) *decompose_ms = b->decompose_ms;
    if (densest_ms) *densest_ms = b->densest_ms;
    return HG_OK;
}

int hg_batch_info(hg_batch* b, int32_t* count, int64_t* items, int64_t* directed_entries, int64_t* device_bytes) {
    if (!b) return fail(HG_ERR_ARG, "batch is NULL");
    if (count) *count = b->count;
    if (items) *items = int64_t(b->items.size());
    if (directed_entries) *directed_entries = int64_t(b->totD);
    if (device_bytes) *device_bytes = b->device_bytes;
    return HG_OK;
}

int hg_batch_destroy(hg_batch* b) {
    if (!b) return fail(HG_ERR_ARG, "batch is NULL");
    if (b->live) (void)wait_live_impl(b);  // its items may still be on the GPU
    release(b);
    return HG_OK;
}

}  // extern "C"
