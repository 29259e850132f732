// hipgrid_batch_kernels.h -- batched FirmCore: many graphs, one persistent launch.
//
// Work model.  A batch holds many independent multilayer graphs resident in
// HBM.  Every (graph, lmd) pair is one work item; the items sit in a device
// queue ordered by estimated cost, largest first, and a fixed grid of
// workgroups pulls them with one atomic per item.  A workgroup runs its
// item's whole decomposition by itself: level 0, then for k = 0, 1, ... the
// cascade of level k and the scan of level k+1, until no vertex is alive.
// Nothing is shared between workgroups except the item counter, so a level
// needs no grid-wide synchronisation and a launch runs all items of all
// graphs to completion.  This is the reference's per-cell PeelSync
// (CoreParallel/FCSyncLeft.cpp:40-131) applied to one (graph, lmd) chain
// (FCSyncLeft.cpp:155-170 walks the k-chain of one lmd), with the chains of
// all graphs and all lmd run as concurrent tasks the way ParGridPlus runs
// its lmd-chains as tasks (FCSyncLeft.cpp:371-409).
//
// Per item, in the item's own slices of the state pools:
//   deg   uint32 [n][L]  residual layer degrees (level 0 pulls them)
//   cnt   int32 [n]      layers with degree >= the threshold
//   idx   int32 [n]      core index (internal ids; -1 while alive)
//   list  uint32 [n]     alive list, compacted in place by every scan
//   queue uint32 [n]     removal FIFO (each vertex enters once per item)
//   bits                 liveness bitmap: in LDS when n <= kLdsBits, else
//                        the item's words in the bitmap pool
// The same rule as the single-graph engine (FCSyncLeft.cpp:65-108): a scan
// removes a vertex with fewer than lmd layers of degree >= k+1; in the
// cascade the decrement taking deg[u][l] from k+1 to k decrements cnt[u],
// and the one taking cnt[u] from lmd to lmd-1 removes u (every vertex that
// is not alive holds cnt < lmd, so that decrement is unique).
//
// Vertices of each graph are relabelled by total degree, descending, at
// upload.  Level 0 of the items of graphs with at most kPullMaxL items comes
// from grid-wide pull kernels that run before the peel (k_batch_pull_*); a
// workgroup pulls the rows of the other graphs' items itself, the heavy
// head (more than kWaveVertex entries) a wavefront per vertex and the rest
// a thread per vertex.  The degree rows of an item's heaviest vertices are
// kept in LDS for the whole chain.  Results are written back in the
// caller's ids at the end of each item.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hipgrid_kernels.h"

namespace hgb {

using hg::div_L;
using hg::kWave;
using hg::lane_id;
using hg::ld_agent;
using hg::sync_stores;
using hg::wave_excl_scan;

constexpr uint32_t kLdsBits = 1u << 18;        // vertices whose liveness bitmap fits in LDS (32 KB)
constexpr uint32_t kLdsWords = kLdsBits / 32;
constexpr uint32_t kWaveVertex = 64;          // level 0: vertices with more entries get a wavefront

struct BGraph {
    uint64_t eb;  // first entry of the graph in adj / adjz
    uint64_t vb;  // first vertex of the graph in z / perm
    uint64_t ob;  // first offset of the graph in off (n + 1 entries, local to eb)
    uint64_t db;  // first word of the graph's deg0 rows [n][L]
    uint32_t n, L, lmagic, H;  // H: vertices with more than kWaveVertex entries (ids < H)
    uint32_t item0, nit;       // the graph's items: ids item0 .. item0 + nit - 1, lmd = 1 .. nit
    uint32_t pulled;           // 1: k_batch_pull writes level 0 of every item (nit <= kPullMaxL)
    uint32_t Hp;               // vertices with more than kPullHeavy entries (ids < Hp)
    uint32_t hub16;            // 1: every layer degree < 2^16, LDS hub rows hold two degrees a word
    uint32_t pad;
};

struct BItem {
    uint32_t g, lmd;
    uint64_t degb;  // deg rows [n][L] in the degree pool
    uint64_t sb;    // n-sized slices: cnt, idx, list, queue pools
    uint64_t bitb;  // bitmap words in the bitmap pool (used when n > kLdsBits)
    uint64_t outb;  // core-index row [n] (caller's ids) in out
    uint64_t histb; // edge histogram [L][kcap] in the histogram pool
    uint32_t kcap;  // histogram width (> every core index of the graph)
    uint32_t id;    // item id (slot of its kmax)
};

// Exclusive prefix sum over a block of B threads; s_tmp needs B/64 + 1 words.
template <int B>
__device__ __forceinline__ uint32_t bscan(uint32_t x, uint32_t* total, uint32_t* s_tmp) {
    constexpr int W = B / kWave;
    const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
    uint32_t wt;
    const uint32_t wex = wave_excl_scan(x, &wt);
    if (lane == 0) s_tmp[wid] = wt;
    __syncthreads();
    if (wid == 0) {
        const uint32_t v = lane < uint32_t(W) ? s_tmp[lane] : 0u;
        uint32_t t;
        const uint32_t e = wave_excl_scan(v, &t);
        if (lane < uint32_t(W)) s_tmp[lane] = e;
        if (lane == 0) s_tmp[W] = t;
    }
    __syncthreads();
    *total = s_tmp[W];
    const uint32_t r = wex + s_tmp[wid];
    __syncthreads();
    return r;
}

// Phase boundaries inside an item.  The item's degree, count, list and
// queue words are written with plain stores by one wave and then read with
// device-scope loads or updated with device-scope atomics by other waves of
// the same workgroup.  All of them run on one CU and reach memory through
// that CU's L2, so waiting until the stores are complete and then meeting
// at the barrier (hg::sync_stores) is enough.  A device-scope fence
// (__threadfence) would in addition write back the whole L2 of the XCD
// (buffer_wbl2) at every boundary: measured on the 256-graph batch, those
// write-backs took 7 % of the launch.
__device__ __forceinline__ void sync_global() {
    sync_stores();
}


// Degree words of the item's heaviest vertices (internal ids < T) live in
// LDS for the whole chain: their decrements are LDS atomics instead of
// device-scope ones.  Ids are sorted by total degree, so the table holds the
// rows that take the most decrements.  The table fills what LDS is left:
// 118 KB for 1024-thread workgroups (one per CU), 32 KB for 512 and 16 KB
// for 256 (two and three per CU), plus the part of the 32 KB bitmap region
// that the item's bitmap leaves free.  When every layer degree of the graph
// is below 2^16 a table word holds two degrees (entry e in half e & 1 of
// word e >> 1; a live decrement never takes a degree below zero, so no
// borrow crosses the halves) and T doubles.  Measured on graphs of
// 90k-180k vertices, a 96 KB table takes 17-32 % of all decrements with one
// degree a word and 24-40 % with two, most of them for graphs of a few ten
// thousand vertices.
template <int B>
constexpr uint32_t hub_words() {
    return B >= 1024 ? 30208u : B >= 512 ? 8192u : 4096u;
}

template <int B>
struct PeelShared {
    // the liveness bitmap (its first ceil(n / 32) words, when n <= kLdsBits),
    // then the hub table in the rest
    uint32_t pool[kLdsWords + hub_words<B>()];
    uint32_t pre[B];
    uint32_t start[B];
    uint32_t tmp[B / kWave + 1];
    uint32_t item;
    uint32_t listlen;
    uint32_t keep;
    uint32_t qhead;
    uint32_t qtail;
    uint32_t nalive;
    uint32_t kmax;
};

struct ItemView {
    const uint32_t* off;   // local offsets [n+1]
    const uint32_t* adj;   // entries u*L + l
    const uint8_t* adjz;   // z(u) per entry
    const uint8_t* z;      // z per vertex
    const uint32_t* deg0;  // [n][L]
    const uint32_t* perm;  // internal id -> caller's id
    uint32_t* deg;
    int32_t* cnt;
    int32_t* idx;
    uint32_t* list;
    uint32_t* queue;
    uint32_t* bits;        // generic pointer: LDS or global
    bool lds_bits;
    bool pulled;           // level-0 rows and cnt written by k_batch_pull
    uint32_t n, L, lmd, magic, H;
    uint32_t T, TL;        // vertices < T have their degree rows in LDS (entries < TL = T * L)
    uint32_t hub0;         // first pool word of the hub table
    bool hub16;            // two degrees a table word
};

// LDS hub table: degree of entry e < TL (reads), decrement returning the old degree.
template <int B>
__device__ __forceinline__ uint32_t hub_get(const ItemView& v, const PeelShared<B>* sh, uint32_t e) {
    const uint32_t* hub = sh->pool + v.hub0;
    return v.hub16 ? (hub[e >> 1] >> (16 * (e & 1))) & 0xFFFFu : hub[e];
}
template <int B>
__device__ __forceinline__ uint32_t hub_dec(const ItemView& v, PeelShared<B>* sh, uint32_t e) {
    uint32_t* hub = sh->pool + v.hub0;
    if (!v.hub16) return atomicSub(&hub[e], 1u);
    const uint32_t sh16 = 16 * (e & 1);
    return (atomicSub(&hub[e >> 1], 1u << sh16) >> sh16) & 0xFFFFu;
}

__device__ __forceinline__ bool bit_get(const ItemView& v, uint32_t u) {
    const uint32_t w = v.lds_bits ? v.bits[u >> 5] : __hip_atomic_load(&v.bits[u >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return (w >> (u & 31)) & 1u;
}
__device__ __forceinline__ void bit_clear(const ItemView& v, uint32_t u) { atomicAnd(&v.bits[u >> 5], ~(1u << (u & 31))); }

// Removes u at level k: index, liveness, FIFO.  Called by the one thread
// whose decrement (or scan) decided the removal.
template <int B>
__device__ __forceinline__ void remove_vertex(const ItemView& v, PeelShared<B>* sh, uint32_t u, int32_t k) {
    v.idx[u] = k;
    bit_clear(v, u);
    const uint32_t q = atomicAdd(&sh->qtail, 1u);
    if (q < v.n) v.queue[q] = u;  // each vertex enters once: q < n always holds
    atomicSub(&sh->nalive, 1u);
}

// Level 0.  Rows: deg[v][l] = #{w in N_l(v) : z(w) >= lmd} for vertices with
// z(v) >= lmd (the others leave at once and no row counts them); cnt = the
// number of non-empty rows.  Then a vertex is alive iff z(v) >= lmd and
// cnt >= lmd; the counted ones that are not alive go to the FIFO.
template <int B>
__device__ void peel_level0(const ItemView& v, PeelShared<B>* sh) {
    const uint32_t n = v.n, L = v.L, lmd = v.lmd;
    const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
    constexpr uint32_t W = B / kWave;
    const uint32_t H = v.pulled ? 0u : v.H;
    const uint32_t n_pull = v.pulled ? 0u : n;
    // heavy head: a wavefront per vertex
    for (uint32_t u = wid; u < H; u += W) {
        if (v.z[u] < lmd) {
            if (lane == 0) v.cnt[u] = 0;
            continue;
        }
        uint32_t pos = v.off[u], c1 = 0;
        for (uint32_t l = 0; l < L; ++l) {
            const uint32_t d = v.deg0[uint64_t(u) * L + l];
            uint32_t c = 0;
            for (uint32_t t = lane; t < d; t += kWave) c += v.adjz[pos + t] >= lmd;
            for (int o = kWave / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, kWave);
            pos += d;
            if (lane == 0) v.deg[uint64_t(u) * L + l] = c;
            c1 += c >= 1u;
        }
        if (lane == 0) v.cnt[u] = int32_t(c1);
    }
    // the rest: a thread per vertex
    for (uint32_t u = H + threadIdx.x; u < n_pull; u += B) {
        uint32_t c1 = 0;
        if (v.z[u] >= lmd) {
            uint32_t pos = v.off[u];
            for (uint32_t l = 0; l < L; ++l) {
                const uint32_t d = v.deg0[uint64_t(u) * L + l];
                uint32_t c = 0;
                for (uint32_t t = 0; t < d; ++t) c += v.adjz[pos + t] >= lmd;
                pos += d;
                v.deg[uint64_t(u) * L + l] = c;
                c1 += c >= 1u;
            }
        }
        v.cnt[u] = int32_t(c1);
    }
    sync_global();
    // classification, coalesced
    for (uint32_t base = 0; base < n; base += B) {
        const uint32_t u = base + threadIdx.x;
        const bool valid = u < n;
        const bool counted = valid && v.z[u] >= lmd;
        const int32_t c1 = valid ? __hip_atomic_load(&v.cnt[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
        const bool alive = counted && c1 >= int32_t(lmd);
        if (valid) v.idx[u] = alive ? -1 : 0;
        // bitmap words: this wavefront's 64 ids are one aligned run
        const uint64_t am = __ballot(alive);
        const uint32_t run0 = base + wid * kWave;
        if (lane == 0 && run0 < n) {
            v.bits[run0 >> 5] = uint32_t(am);
            if (run0 + 32 < n) v.bits[(run0 >> 5) + 1] = uint32_t(am >> 32);
        }
        if (am) {
            uint32_t kb = 0;
            if (lane == 0) kb = atomicAdd(&sh->listlen, uint32_t(__popcll(am)));
            kb = __shfl(kb, 0, kWave);
            if (alive) v.list[kb + __builtin_amdgcn_mbcnt_hi(uint32_t(am >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(am), 0u))] = u;
        }
        const uint64_t qm = __ballot(counted && !alive);
        if (qm) {
            uint32_t qb = 0;
            if (lane == 0) qb = atomicAdd(&sh->qtail, uint32_t(__popcll(qm)));
            qb = __shfl(qb, 0, kWave);
            if (counted && !alive)
                v.queue[qb + __builtin_amdgcn_mbcnt_hi(uint32_t(qm >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(qm), 0u))] = u;
        }
    }
    // the hub rows move to LDS (rows of vertices with z < lmd are never read)
    if (v.hub16) {
        // A word can pair the last row of one vertex with the first of the
        // next (odd L), and the rows of a vertex with z < lmd hold whatever
        // the pool held before: each half is cut to 16 bits so that such a
        // word never spills into its valid neighbour.
        for (uint32_t w = threadIdx.x; 2 * w < v.TL; w += B) {
            const uint32_t lo = ld_agent(&v.deg[2 * w]) & 0xFFFFu;
            const uint32_t hi = 2 * w + 1 < v.TL ? ld_agent(&v.deg[2 * w + 1]) & 0xFFFFu : 0u;
            sh->pool[v.hub0 + w] = lo | (hi << 16);
        }
    } else {
        for (uint32_t e = threadIdx.x; e < v.TL; e += B) sh->pool[v.hub0 + e] = ld_agent(&v.deg[e]);
    }
    sync_global();
    if (threadIdx.x == 0) sh->nalive = sh->listlen;
    __syncthreads();
}

// Cascade of level k: expands the FIFO from qhead until it is empty, B
// vertices per round, their entries spread over the block (U per thread in
// flight).
template <int B, int U>
__device__ void peel_cascade(const ItemView& v, PeelShared<B>* sh, int32_t k) {
    const uint32_t thr = uint32_t(k) + 1u, lmd = v.lmd;
    for (;;) {
        sync_stores();  // FIFO entries, liveness bits and degrees of the last round are settled
        const uint32_t qh = sh->qhead, qt = sh->qtail < v.n ? sh->qtail : v.n;
        if (qh >= qt) break;
        const uint32_t nb = qt - qh < uint32_t(B) ? qt - qh : uint32_t(B);
        uint32_t start = 0, len = 0;
        if (threadIdx.x < nb) {
            const uint32_t u = ld_agent(&v.queue[qh + threadIdx.x]);
            start = v.off[u];
            len = v.off[u + 1] - start;
        }
        uint32_t total;
        const uint32_t pre = bscan<B>(len, &total, sh->tmp);
        sh->pre[threadIdx.x] = pre;
        sh->start[threadIdx.x] = start;
        __syncthreads();
        for (uint32_t j0 = 0; j0 < total; j0 += uint32_t(B) * U) {
            uint32_t e[U], u[U], old[U];
            bool live[U];
#pragma unroll
            for (int r = 0; r < U; ++r) {
                const uint32_t j = j0 + uint32_t(r) * B + threadIdx.x;
                live[r] = j < total;
                e[r] = 0;
                if (live[r]) {
                    uint32_t lo = 0, hi = nb - 1;  // largest p with pre[p] <= j
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi + 1) >> 1;
                        if (sh->pre[mid] <= j) lo = mid;
                        else hi = mid - 1;
                    }
                    const uint32_t p = sh->start[lo] + (j - sh->pre[lo]);
                    e[r] = v.adj[p];
                    live[r] = v.adjz[p] >= lmd;
                }
            }
#pragma unroll
            for (int r = 0; r < U; ++r) {
                u[r] = div_L(e[r], v.L, v.magic);
                if (live[r]) live[r] = bit_get(v, u[r]);
            }
#pragma unroll
            for (int r = 0; r < U; ++r)
                old[r] = !live[r] ? 0u : e[r] < v.TL ? hub_dec<B>(v, sh, e[r]) : atomicSub(&v.deg[e[r]], 1u);
            int32_t oc[U];
#pragma unroll
            for (int r = 0; r < U; ++r) oc[r] = (live[r] && old[r] == thr) ? atomicSub(&v.cnt[u[r]], 1) : 0;
#pragma unroll
            for (int r = 0; r < U; ++r)
                if (live[r] && old[r] == thr && oc[r] == int32_t(lmd)) remove_vertex<B>(v, sh, u[r], k);
        }
        __syncthreads();
        if (threadIdx.x == 0) sh->qhead = qh + nb;
    }
    __syncthreads();
}

// #{l < L : row[l] >= thr} for a degree row in global memory, two words per
// device-scope load where the row is 8-byte aligned (half the requests of a
// word at a time; each 64-bit load sees two words that atomics update).
__device__ __forceinline__ uint32_t row_count(const uint32_t* row, uint32_t L, uint32_t thr) {
    uint32_t c = 0, l = 0;
    if ((reinterpret_cast<uintptr_t>(row) & 7u) != 0 && L > 0) {
        c += ld_agent(&row[0]) >= thr;
        l = 1;
    }
    for (; l + 2 <= L; l += 2) {
        const unsigned long long w = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(row + l),
                                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        c += uint32_t(uint32_t(w) >= thr) + uint32_t(uint32_t(w >> 32) >= thr);
    }
    if (l < L) c += ld_agent(&row[l]) >= thr;
    return c;
}

// Scan of level k >= 1 over the alive list, compacted in place.
template <int B>
__device__ void peel_scan(const ItemView& v, PeelShared<B>* sh, int32_t k) {
    const uint32_t thr = uint32_t(k) + 1u, L = v.L, lmd = v.lmd;
    const uint32_t len = sh->listlen;
    if (threadIdx.x == 0) sh->keep = 0;
    __syncthreads();
    for (uint32_t base = 0; base < len; base += B) {
        const uint32_t t = base + threadIdx.x;
        uint32_t u = 0;
        bool live = false;
        if (t < len) {
            u = ld_agent(&v.list[t]);
            live = bit_get(v, u);
        }
        uint32_t c = 0;
        if (live) {
            if (u < v.T)
                for (uint32_t l = 0; l < L; ++l) c += hub_get<B>(v, sh, u * L + l) >= thr;
            else
                c = row_count(v.deg + uint64_t(u) * L, L, thr);
        }
        const bool rem = live && c < lmd;
        const bool keep = live && !rem;
        if (live) v.cnt[u] = int32_t(c);
        if (rem) remove_vertex<B>(v, sh, u, k);
        __syncthreads();  // the chunk is read before compaction writes into it
        const uint64_t km = __ballot(keep);
        if (km) {
            uint32_t kb = 0;
            if (lane_id() == 0) kb = atomicAdd(&sh->keep, uint32_t(__popcll(km)));
            kb = __shfl(kb, 0, kWave);
            if (keep) v.list[kb + __builtin_amdgcn_mbcnt_hi(uint32_t(km >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(km), 0u))] = u;
        }
    }
    sync_global();  // the cnt stores above feed the cascade's atomics
    if (threadIdx.x == 0) sh->listlen = sh->keep;
    __syncthreads();
}

// ------------------------------------------------------- batched level 0
// Level 0 of every item of a graph in one pass over its adjacency, for all
// the graphs of the batch in one grid-wide launch ahead of the peel: for
// each vertex v and layer l the counts
//     deg_lmd[v][l] = #{w in N_l(v) : z(w) >= lmd},   lmd = 1 .. nit,
// come out of the byte-parallel count of the z stream (hg::count_word), and
// each item's row and cnt (non-empty rows) are written to its pool slices.
// Rows are written only for vertices with z(v) >= lmd (the others leave at
// level 0 and no live entry points at them); cnt = 0 for the others.  A
// workgroup running item (g, lmd) then only classifies.  Graphs with more
// than kPullMaxL items keep the per-item pull in peel_level0.
constexpr uint32_t kPullBlock = 256;
constexpr uint32_t kPullHeavyTile = 2 * kPullBlock / kWave;  // a wavefront per vertex, two per wavefront
// Light tiles: a thread per (vertex, layer) row, at most kPullBlock rows and
// kPullStageBytes adjacency bytes (staged in LDS) per tile.
constexpr uint32_t kPullStageBytes = 8192;

struct PullTile {
    uint32_t g;      // graph
    uint32_t v0;     // first vertex (internal id)
    uint32_t count;  // vertices
    uint32_t heavy;  // 1: hubs (more than kPullStageBytes entries), a wavefront per vertex
};

struct PullItems {
    uint64_t degb[hg::kPullMaxL];
    uint64_t sb[hg::kPullMaxL];
};

// Hubs (vertices with more than kPullStageBytes entries, the first ids): a
// wavefront per vertex, lanes strided over the 16-byte words of each layer
// segment, counts reduced across the wavefront.
template <uint32_t NJ>
__global__ void __launch_bounds__(kPullBlock) k_batch_pull_heavy(const BGraph* graphs, const BItem* items_by_id,
                                                                const PullTile* tiles, uint32_t ntiles,
                                                                const uint32_t* off, const uint8_t* adjz,
                                                                const uint8_t* z, const uint32_t* deg0,
                                                                uint32_t* degpool, int32_t* cntpool) {
    __shared__ PullItems s_it;
    const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const PullTile tile = tiles[t];
        const BGraph gr = graphs[tile.g];
        __syncthreads();  // s_it of the previous tile is no longer read
        if (threadIdx.x < gr.nit) {
            const BItem it = items_by_id[gr.item0 + threadIdx.x];
            s_it.degb[threadIdx.x] = it.degb;
            s_it.sb[threadIdx.x] = it.sb;
        }
        __syncthreads();
        const uint32_t L = gr.L;
        for (uint32_t i = wid; i < tile.count; i += kPullBlock / kWave) {
            const uint32_t v = tile.v0 + i;
            const uint32_t zv = z[gr.vb + v];
            uint32_t mine = 0;  // lane j: cnt of lmd = j + 1
            if (zv > 0) {
                uint64_t pos = gr.eb + off[gr.ob + v];
                for (uint32_t l = 0; l < L; ++l) {
                    const uint32_t d = deg0[gr.db + uint64_t(v) * L + l];
                    uint32_t c[NJ];
#pragma unroll
                    for (uint32_t j = 0; j < NJ; ++j) c[j] = 0;
                    hg::pull_count_vec<NJ>(adjz, pos, pos + d, lane, kWave, c);
                    pos += d;
#pragma unroll
                    for (uint32_t j = 0; j < NJ; ++j)
                        for (int o = kWave / 2; o > 0; o >>= 1) c[j] += __shfl_xor(c[j], o, kWave);
                    uint32_t cj = 0;
#pragma unroll
                    for (uint32_t j = 0; j < NJ; ++j)
                        if (lane == j) cj = c[j];
                    if (lane < gr.nit && lane < zv) {
                        degpool[s_it.degb[lane] + uint64_t(v) * L + l] = cj;
                        mine += cj != 0u;
                    }
                }
            }
            if (lane < gr.nit) cntpool[s_it.sb[lane] + v] = lane < zv ? int32_t(mine) : 0;
        }
    }
}

// The rest: a thread per (vertex, layer) row.  The tile's adjacency bytes
// are staged in LDS with coalesced word loads and its deg0 rows read
// coalesced; each thread counts its segment from LDS, the item rows are
// written coalesced (row r of the tile is word v0 * L + r of every item's
// rows), and the non-empty row counts are summed per vertex in LDS.
// W = 8: vertices with at most kPullHeavy entries, counts packed four to a
// word; W = 16: vertices with at most kPullStageBytes entries, two to a word.
template <uint32_t NJ, uint32_t W>
__global__ void __launch_bounds__(kPullBlock) k_batch_pull_light(const BGraph* graphs, const BItem* items_by_id,
                                                                const PullTile* tiles, uint32_t ntiles,
                                                                const uint32_t* off, const uint8_t* adjz,
                                                                const uint8_t* z, const uint32_t* deg0,
                                                                uint32_t* degpool, int32_t* cntpool) {
    __shared__ PullItems s_it;
    __shared__ uint32_t s_stage[kPullStageBytes / 4 + 2];
    __shared__ uint32_t s_cl[kPullBlock * (NJ / 4)];  // per vertex of the tile: packed non-empty row counts
    __shared__ uint32_t s_tmp[kPullBlock / kWave + 1];
    const uint8_t* stage = reinterpret_cast<const uint8_t*>(s_stage);
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const PullTile tile = tiles[t];
        const BGraph gr = graphs[tile.g];
        const uint32_t L = gr.L, rows = tile.count * L;
        __syncthreads();  // LDS of the previous tile is no longer read
        if (threadIdx.x < gr.nit) {
            const BItem it = items_by_id[gr.item0 + threadIdx.x];
            s_it.degb[threadIdx.x] = it.degb;
            s_it.sb[threadIdx.x] = it.sb;
        }
        for (uint32_t i = threadIdx.x; i < tile.count * (NJ / 4); i += kPullBlock) s_cl[i] = 0;
        // the tile's bytes [beg, end) of adjz, from the aligned word below beg
        const uint64_t beg = gr.eb + off[gr.ob + tile.v0], end = gr.eb + off[gr.ob + tile.v0 + tile.count];
        const uint64_t w0 = beg >> 2;
        const uint32_t nw = uint32_t(((end + 3) >> 2) - w0);  // <= kPullStageBytes / 4 + 1
        const uint32_t* zw = reinterpret_cast<const uint32_t*>(adjz);
        for (uint32_t i = threadIdx.x; i < nw; i += kPullBlock) s_stage[i] = zw[w0 + i];
        // this thread's row: its length and start (block scan of the lengths)
        const uint32_t r = threadIdx.x;
        const uint32_t d = r < rows ? deg0[gr.db + uint64_t(tile.v0) * L + r] : 0u;
        uint32_t total;
        const uint32_t pre = bscan<kPullBlock>(d, &total, s_tmp);  // ends with a barrier
        if (r < rows) {
            const uint32_t vl = r / L;
            const uint32_t v = tile.v0 + vl;
            const uint32_t zv = z[gr.vb + v];
            constexpr uint32_t P = 32 / W;  // counts a word
            constexpr uint32_t M = (1u << W) - 1u;
            constexpr uint32_t ONES = W == 8 ? 0x01010101u : 0x00010001u;
            uint32_t cw[NJ / P];
#pragma unroll
            for (uint32_t q = 0; q < NJ / P; ++q) cw[q] = 0;
            const uint32_t s0 = uint32_t(beg - (w0 << 2)) + pre;
            for (uint32_t k = 0; k < d; ++k) {
                const int32_t zb = stage[s0 + k];  // entries with z > j for j < zb
#pragma unroll
                for (uint32_t q = 0; q < NJ / P; ++q) {
                    const int32_t m = zb - int32_t(P * q);  // counts of word q that this entry raises
                    cw[q] += m >= int32_t(P) ? ONES : m <= 0 ? 0u : ONES & ((1u << (W * m)) - 1u);
                }
            }
            const uint64_t row = uint64_t(tile.v0) * L + r;
#pragma unroll
            for (uint32_t j = 0; j < NJ; ++j)
                if (j < gr.nit && j < zv) degpool[s_it.degb[j] + row] = (cw[j / P] >> (W * (j % P))) & M;
            // non-empty rows: byte j of word q is 1 when count j is non-zero
#pragma unroll
            for (uint32_t q = 0; q < NJ / 4; ++q) {
                uint32_t nz = 0;
#pragma unroll
                for (uint32_t b = 0; b < 4; ++b) {
                    const uint32_t j = 4 * q + b;
                    nz |= uint32_t(((cw[j / P] >> (W * (j % P))) & M) != 0u) << (8 * b);
                }
                if (nz) atomicAdd(&s_cl[vl * (NJ / 4) + q], nz);
            }
        }
        __syncthreads();
        if (threadIdx.x < tile.count) {
            const uint32_t v = tile.v0 + threadIdx.x;
            const uint32_t zv = z[gr.vb + v];
#pragma unroll
            for (uint32_t j = 0; j < NJ; ++j)
                if (j < gr.nit)
                    cntpool[s_it.sb[j] + v] =
                        j < zv ? int32_t((s_cl[threadIdx.x * (NJ / 4) + (j >> 2)] >> (8 * (j & 3))) & 0xFFu) : 0;
        }
    }
}

// Resets the item counter (a kernel, so that a captured graph holds no
// memset node).
__global__ void k_batch_reset(uint32_t* head) {
    if (threadIdx.x == 0) head[0] = 0;
}

// Pool base pointers of a batch (each item's slices are offsets into them).
struct BPools {
    const uint32_t* off;
    const uint32_t* adj;
    const uint8_t* adjz;
    const uint8_t* z;
    const uint32_t* deg0;
    const uint32_t* perm;
    uint32_t* degpool;
    int32_t* cntpool;
    int32_t* idxpool;
    uint32_t* listpool;
    uint32_t* qpool;
    uint32_t* bitpool;
    int32_t* out;
    int32_t* kmax_out;
    unsigned long long* item_clock;  // diagnostic (nullptr: off): [item id][2] start, end (wall clock)
};

// One item's whole k-chain (level 0, then cascade / scan per level) and its
// results (core-index row in the caller's ids, kmax).  All threads call;
// sh is free on entry and on return.
template <int B, int U>
__device__ void run_item(PeelShared<B>* sh, const BItem& item, const BGraph& gr, const BPools& P, bool pulled) {
    if (P.item_clock && threadIdx.x == 0) P.item_clock[2 * item.id] = wall_clock64();
    ItemView v;
    v.off = P.off + gr.ob;
    v.adj = P.adj + gr.eb;
    v.adjz = P.adjz + gr.eb;
    v.z = P.z + gr.vb;
    v.deg0 = P.deg0 + gr.db;
    v.perm = P.perm + gr.vb;
    v.deg = P.degpool + item.degb;
    v.cnt = P.cntpool + item.sb;
    v.idx = P.idxpool + item.sb;
    v.list = P.listpool + item.sb;
    v.queue = P.qpool + item.sb;
    v.lds_bits = gr.n <= kLdsBits;
    v.bits = v.lds_bits ? sh->pool : P.bitpool + item.bitb;
    v.n = gr.n;
    v.L = gr.L;
    v.lmd = item.lmd;
    v.magic = gr.lmagic;
    v.H = gr.H;
    v.pulled = pulled;
    v.hub16 = gr.hub16 != 0;
    {
        // the bitmap words in LDS (rounded to 16 bytes), the table after them
        v.hub0 = v.lds_bits ? ((gr.n + 31) / 32 + 3) & ~3u : 0u;
        const uint32_t cap = (v.hub16 ? 2u : 1u) * (kLdsWords + hub_words<B>() - v.hub0) / gr.L;
        v.T = gr.n < cap ? gr.n : cap;
    }
    v.TL = v.T * gr.L;
    if (threadIdx.x == 0) {
        sh->listlen = 0;
        sh->qhead = 0;
        sh->qtail = 0;
        sh->kmax = 0;
    }
    __syncthreads();
    peel_level0<B>(v, sh);
    // every vertex is removed by level max-degree + 1 at the latest
    // (item.kcap = max layer degree + 2), which also bounds the loop
    int32_t k = 0;
    for (;;) {
        peel_cascade<B, U>(v, sh, k);
        if (sh->nalive == 0 || uint32_t(k) + 1 >= item.kcap) break;
        ++k;
        peel_scan<B>(v, sh, k);
    }
    // every vertex got its index at level <= k; k is the largest index
    sync_global();
    for (uint32_t u = threadIdx.x; u < v.n; u += B)
        P.out[item.outb + v.perm[u]] = int32_t(ld_agent(reinterpret_cast<const uint32_t*>(&v.idx[u])));
    if (threadIdx.x == 0) P.kmax_out[item.id] = k;
    if (P.item_clock && threadIdx.x == 0) P.item_clock[2 * item.id + 1] = wall_clock64();
    __syncthreads();
}

template <int B, int U>
__global__ void __launch_bounds__(B) k_batch_peel(const BGraph* graphs, const BItem* items, uint32_t nitems,
                                                  uint32_t* head, const uint32_t* off, const uint32_t* adj,
                                                  const uint8_t* adjz, const uint8_t* z, const uint32_t* deg0,
                                                  const uint32_t* perm, uint32_t* degpool, int32_t* cntpool,
                                                  int32_t* idxpool, uint32_t* listpool, uint32_t* qpool,
                                                  uint32_t* bitpool, int32_t* out, int32_t* kmax_out,
                                                  unsigned long long* item_clock) {
    __shared__ PeelShared<B> sh;
    const BPools P{off,     adj,      adjz,  z,   deg0,     perm,      degpool,   cntpool,
                   idxpool, listpool, qpool, bitpool, out, kmax_out, item_clock};
    for (;;) {
        if (threadIdx.x == 0) sh.item = atomicAdd(head, 1u);
        __syncthreads();
        const uint32_t it = sh.item;
        if (it >= nitems) break;
        const BItem item = items[it];
        const BGraph gr = graphs[item.g];
        run_item<B, U>(&sh, item, gr, P, gr.pulled != 0);
    }
}

// Edge histogram of the listed items (a block per item, round robin):
// hist[l][m] += 1 for every entry v -> (u, l) with m = min(idx v, idx u) >= 1.
template <int B>
__global__ void __launch_bounds__(B) k_batch_edge_hist(const BGraph* graphs, const BItem* items, uint32_t nsel,
                                                       const uint32_t* off, const uint32_t* adj,
                                                       const int32_t* idxpool, uint32_t* histpool) {
    __shared__ uint32_t s_pre[B];
    __shared__ uint32_t s_start[B];
    __shared__ int32_t s_iv[B];
    __shared__ uint32_t s_tmp[B / kWave + 1];
    for (uint32_t si = blockIdx.x; si < nsel; si += gridDim.x) {
        const BItem item = items[si];
        const BGraph gr = graphs[item.g];
        const uint32_t* goff = off + gr.ob;
        const uint32_t* gadj = adj + gr.eb;
        const int32_t* idx = idxpool + item.sb;
        uint32_t* hist = histpool + item.histb;
        for (uint32_t v0 = 0; v0 < gr.n; v0 += B) {
            const uint32_t vv = v0 + threadIdx.x;
            uint32_t start = 0, len = 0;
            int32_t iv = 0;
            if (vv < gr.n) {
                iv = idx[vv];
                if (iv >= 1) {
                    start = goff[vv];
                    len = goff[vv + 1] - start;
                }
            }
            uint32_t total;
            const uint32_t pre = bscan<B>(len, &total, s_tmp);
            s_pre[threadIdx.x] = pre;
            s_start[threadIdx.x] = start;
            s_iv[threadIdx.x] = iv;
            __syncthreads();
            for (uint32_t j = threadIdx.x; j < total; j += B) {
                uint32_t lo = 0, hi = B - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (s_pre[mid] <= j) lo = mid;
                    else hi = mid - 1;
                }
                const uint32_t e = gadj[s_start[lo] + (j - s_pre[lo])];
                const uint32_t u = div_L(e, gr.L, gr.lmagic);
                const uint32_t l = e - u * gr.L;
                const int32_t iu = idx[u];
                if (iu >= 1) {
                    const uint32_t m = uint32_t(iu < s_iv[lo] ? iu : s_iv[lo]);
                    atomicAdd(&hist[uint64_t(l) * item.kcap + m], 1u);
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace hgb

namespace hgb {

// ------------------------------------------------------------- live queue
// A device work queue shared by batches (hg_live): the host appends items
// of any batch to a ring of slots and advances `tail`; every workgroup of
// every running k_live_peel launch claims the next slot with a CAS on
// `head` whenever it has finished an item, so items submitted while earlier
// ones run are taken by the workgroups of the launches already on the GPU
// as they free up.  A workgroup that finds the queue empty exits at once
// (nothing waits on the GPU); each submission also launches a grid of its
// own, so its items run even when every earlier workgroup has exited.
// head and tail are read and updated only with atomics, which the GPU
// performs past its L2 caches, where the host's copies of tail land.  An
// item's memory may have been another item's before (the allocation cache
// hands freed blocks out again), so a workgroup drops its XCD's cached lines
// before it reads a slot (acquire) and writes its results back (release)
// before it raises the slot's done flag in host memory.
struct alignas(128) LiveCtl {
    unsigned long long head;
    uint32_t pad0[30];
    unsigned long long tail;
    uint32_t pad1[30];
    int32_t error;
    uint32_t pad2[31];
};

struct alignas(256) LiveSlot {
    unsigned long long seq;       // queue position this slot was written for
    unsigned long long instance;  // the launch made for the submission
    BPools P;
    BGraph gr;
    BItem item;
};

enum LiveErr : int32_t { kLiveErrSlot = 1 };

template <int B, int U>
__global__ void __launch_bounds__(B) k_live_peel(LiveCtl* ctl, const LiveSlot* ring, uint32_t ring_cap,
                                                 unsigned long long* flags, unsigned long long instance) {
    __shared__ PeelShared<B> sh;
    __shared__ LiveSlot s_slot;
    __shared__ unsigned long long s_pos;
    for (;;) {
        if (threadIdx.x == 0) {
            unsigned long long pos = ~0ull;
            for (;;) {  // ends: every failed CAS means another workgroup took a slot
                const unsigned long long h = atomicAdd(&ctl->head, 0ull);
                const unsigned long long t = atomicAdd(&ctl->tail, 0ull);
                if (h >= t) break;
                if (atomicCAS(&ctl->head, h, h + 1) == h) {
                    pos = h;
                    break;
                }
            }
            s_pos = pos;
        }
        __syncthreads();
        const unsigned long long pos = s_pos;
        if (pos == ~0ull) break;
        // The host copies a submission's slots before its new tail, on one
        // stream; a slot whose copy has not landed yet is read again (a
        // bounded number of times) before it counts as an error.
        bool ok = false;
        for (uint32_t attempt = 0; attempt < 4096 && !ok; ++attempt) {
            if (threadIdx.x < kWave) __threadfence();  // acquire: the CU's and XCD's cached lines are dropped
            __syncthreads();
            const uint32_t* src = reinterpret_cast<const uint32_t*>(ring + pos % ring_cap);
            uint32_t* dst = reinterpret_cast<uint32_t*>(&s_slot);
            for (uint32_t w = threadIdx.x; w < sizeof(LiveSlot) / 4; w += B) dst[w] = src[w];
            __syncthreads();
            ok = s_slot.seq == pos;
            if (!ok) __builtin_amdgcn_s_sleep(64);
        }
        if (ok) run_item<B, U>(&sh, s_slot.item, s_slot.gr, s_slot.P, false);
        else if (threadIdx.x == 0) atomicCAS(&ctl->error, 0, int32_t(kLiveErrSlot));
        sync_stores();
        if (threadIdx.x < kWave) __threadfence();  // release: results and item state reach memory
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long* f = flags + 2 * (pos % ring_cap);
            __hip_atomic_store(f + 1, instance, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(f, pos + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        __syncthreads();
    }
}

}  // namespace hgb
