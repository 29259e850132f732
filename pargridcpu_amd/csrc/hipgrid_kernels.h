// hipgrid_kernels.h -- CDNA4 (gfx950) kernels of the HipGrid FirmCore engine.
//
// Work model.  A decomposition runs one peel instance per lmd = 1..L over a
// graph resident in HBM.  Per instance and vertex: the L layer degrees in
// the residual graph (deg, AoS [n][L]: one vertex's row is one or two
// 32-byte sectors), cnt = number of layers with degree >= the current
// threshold, and idx = core index (-1 while alive).  Separate compact
// arrays: the random accesses are bound by cache capacity, so the hot
// footprint per vertex is kept small (a 64-byte record per vertex measured
// 1.5x slower on the 1.2B-edge workload).  Level k removes every
// alive vertex whose index is k, turning S(k, lmd) into S(k+1, lmd).  This
// is FCSyncLeft::PeelSync's two-phase rule (CoreParallel/FCSyncLeft.cpp:
// 65-108):
//   scan     a vertex with fewer than lmd layers of degree >= k+1 is removed,
//            the others get cnt;
//   cascade  each entry of a removed vertex decrements deg[u][l] with an
//            atomic; the unique decrement taking it from k+1 to k decrements
//            cnt[u]; the unique decrement taking cnt[u] from lmd to lmd-1
//            removes u.  FCSyncLeft guards the removal with a CAS on valid[u];
//            here every vertex that is not alive holds cnt < lmd, which makes
//            that decrement unique over the whole run, so no CAS is needed.
//
// Scheduling, designed around the cost of same-address device atomics
// (they serialise, ~11 ns each on MI355X, while scattered atomics reach
// ~26 G/s):
//   * Vertices are relabelled by total degree, descending, when the graph is
//     uploaded, so the rows that take most of the random accesses (hubs)
//     form one compact, cache-resident region; results are mapped back to
//     the caller's ids at the end.
//   * The grid is G fixed blocks.  Vertices are dealt to blocks in runs of
//     kRun consecutive ids, round robin (run c goes to block c mod G): each
//     run is one coalesced stretch of rows, and every block gets its share
//     of hubs and of leaves, so the alive lists stay balanced at every level
//     even though ids are sorted by degree.  Block b keeps its alive list
//     compacted in place in its own list slice.
//   * Vertices removed by a scan become adjacency chunks on the step's
//     frontier list (one atomic per block and chunk of vertices); the
//     cascade strides that list over all blocks.  Vertices removed during a
//     cascade go to the discovering block's own FIFO ring (slots reserved
//     with LDS atomics) and are expanded by that block; heavy ones
//     (> heavy entries, default kHeavy) and ring overflow go to the next step's list, so a
//     hub is spread over the whole GPU.  Nothing ever spins.
//   * Liveness is read from a per-instance bitmap (n bits: 100 MB for all
//     eight instances of a 100M-vertex graph, so it stays in the 256 MB
//     Infinity Cache) instead of the 4-byte index: a dead neighbour costs a
//     cached bit, not a random HBM line.  The bit is cleared when the
//     vertex is removed; a stale bit only costs a harmless decrement of a
//     removed vertex (its cnt is already below lmd).
//   * One step = A (scan, or level-0 init; every block first decides, per
//     instance, whether the step starts a new level, continues the current
//     one while deferred chunks remain, or finds the instance done) + B
//     (cascade).
//
// Adjacency is vertex-major: entries of all layers of v are contiguous at
// adj[off[v] .. off[v+1]); an entry is u*L + l, which is also the index of
// deg[u][l].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hg {
// Internal linkage: each translation unit of the library that includes this
// header gets its own copy of the kernels it launches.
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 512;  // threads per block of the level kernels (8 wavefronts)
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kChunk = 64;         // static chunk list granularity (densest pass)
constexpr uint32_t kHeavy = 1024;  // default: vertices with more entries are deferred
constexpr uint32_t kDefer = 1024;     // default entries per deferred chunk
constexpr uint32_t kMinDefer = 256;   // smallest chunk size the lists are sized for
constexpr uint32_t kRun = 64;      // consecutive vertex ids per ownership run

enum Phase : int32_t { kPhaseScan = 0, kPhaseCont = 1, kPhaseDone = 2 };

// Per-instance constant data (one lmd of one graph).
struct Inst {
    const int64_t* off;   // [n+1] merged adjacency offsets
    const uint32_t* adj;  // [D] entries u*L + l
    const uint8_t* adjz;  // [D] z(u) of each entry's target (layers in which u has an edge)
    const uint8_t* adjh;  // [D] or null: min(h, 255), h = the target's lmd-th largest layer degree among neighbours of z >= lmd
    uint32_t* alive;      // [ceil(n/32)] liveness bitmap
    uint32_t* deg;        // [n*L] residual layer degrees
    int32_t* cnt;         // [n]
    int32_t* idx;         // [n] core index, -1 while alive
    uint32_t* list;       // [G*cap] per-block alive lists
    uint32_t* ring;       // [G*cap] per-block removal rings
    uint32_t* blk;        // [G*2] per block: alive list length, spare
    uint64_t* pl0;        // frontier / deferred chunk lists, by step parity
    uint64_t* pl1;
    uint64_t plcap;
    uint32_t n, L, lmd, cap;
    uint32_t lmagic;  // min(floor(2^32 / L), 2^32 - 1)
    uint32_t flags;   // kFlagAggregate: combine decrements in LDS
    uint32_t heavy;   // cascade removals with more entries are deferred to the next step
    uint32_t defer;   // entries per deferred chunk (>= kMinDefer)
    uint32_t tail_heavy;   // the deferral threshold of light steps (tail_chunks chunks or fewer in all)
    uint32_t tail_chunks;
};

__device__ __forceinline__ uint32_t* rec_of(const Inst& in, uint32_t v) { return in.deg + uint64_t(v) * in.L; }
__device__ __forceinline__ int32_t* idx_of(const Inst& in, uint32_t v) { return in.idx + v; }
__device__ __forceinline__ int32_t* cnt_of(const Inst& in, uint32_t v) { return in.cnt + v; }
__device__ __forceinline__ bool is_alive(const Inst& in, uint32_t v) { return (in.alive[v >> 5] >> (v & 31)) & 1u; }
__device__ __forceinline__ void clear_alive(const Inst& in, uint32_t v) {
    atomicAnd(&in.alive[v >> 5], ~(1u << (v & 31)));
}
// Pre-filter of an entry before its target's liveness probe (a streaming
// byte instead of a random bitmap word): the target can be alive at level k
// only if f.z[p] >= f.t.  With the core-index bounds, a target whose bound h
// is below k left at an earlier level, and h >= 1 holds exactly when the
// target has edges in at least lmd layers (the z test); without them, z(u) >=
// lmd.  A saturated bound (255) never filters.
struct EntryFilter {
    const uint8_t* z;
    uint32_t t;
};
__device__ __forceinline__ EntryFilter entry_filter(const Inst& in, int32_t k) {
    if (in.adjh) return {in.adjh, k < 1 ? 1u : k > 255 ? 255u : uint32_t(k)};
    return {in.adjz, in.lmd};
}
constexpr uint32_t kFlagAggregate = 1;
// The instance's rows are a compacted copy (k_cmp_*): every target was alive
// when the copy was made, so the z pre-filter (adjz) does not apply.
constexpr uint32_t kFlagCompact = 2;

// Per-instance state; the atomically updated words sit on lines of their own.
// Step s uses the slots of parity s & 1: its phase and level (written once,
// by block 0 of the step's level kernel), the chunks its scan pushed
// (pfront) and the chunks the previous step's cascade deferred to it
// (pdefer).  One chunk list per parity holds both kinds: a step either
// continues with deferred chunks or scans, never both.
// A scan step of an instance may apply its first removal rounds binned by
// target (k_bin_*, rounds r = 0, 1, ...: round 0 takes the scan's frontier,
// round r > 0 the vertices round r - 1 removed).  bflag[r & 1] says whether
// round r binned, bmode whether any round of the step did and bround the
// last one; rbase is the round's slice of the record buffers; pfent counts
// the frontier's entries (the scan adds them); pbr[r & 1] / pber[r & 1]
// count the chunks and entries of the vertices round r removed, which the
// next round or the cascade takes from the same list instead of the
// frontier.
struct alignas(128) InstState {
    int32_t dphase[2];
    int32_t dk[2];
    int32_t bmode[2];
    int32_t bround[2];
    unsigned long long rbase[2];
    int32_t bflag[2];
    int32_t pad0[18];
    unsigned long long pdefer[2][16];  // one line each
    unsigned long long pfront[2][16];
    unsigned long long alive[16];      // vertices not yet removed (decremented mod 2^64)
    unsigned long long pfent[2][16];
    unsigned long long pbr[2][16];
    unsigned long long pber[2][16];
    unsigned long long brec[16];       // records the binned applies took (cumulative; diagnostics)
    unsigned long long r0cnt[2][16];   // records of binned round r (slot r & 1) in expansion order; the round's apply clears it
};

struct alignas(128) Control {
    int32_t error;   // nonzero: device-side check failed
    int32_t active;  // instances not done at the last level kernel
    int32_t binned;  // (instance, round) pairs applied binned (k_bin_expand, block 0)
    int32_t pad[29];
    unsigned long long dbg[16];  // event counters of the binned kernels (diagnostics)
};

enum Err : int32_t { kErrNone = 0, kErrDefer = 2, kErrCompact = 3, kErrBin = 4 };

// ----------------------------------------------------------------- helpers
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Waves of one block hand each other work through global memory (a block's
// removal ring, a batch item's FIFO and alive list).  A plain block barrier
// waits for LDS operations only; a global store issued before it can still
// be in flight when another wave reads the word after it.  This barrier
// first waits for every memory operation the thread has issued (stores and
// atomics included), so they are complete at the L2 when any wave passes.
__device__ __forceinline__ void sync_stores() {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// Agent-scope load: served by the L2, never by a possibly stale vector L1
// line (words written by other waves within the same kernel).
__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// e / L for e < 2^32 with magic = min(floor(2^32 / L), 2^32 - 1): the
// multiply-high estimate is at most one below the quotient.
__device__ __forceinline__ uint32_t div_L(uint32_t e, uint32_t L, uint32_t magic) {
    uint32_t q = __umulhi(e, magic);
    if (e - q * L >= L) ++q;
    return q;
}

// Exclusive prefix sum of x over the wavefront; *total gets the sum.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t x, uint32_t* total) {
    const uint32_t lane = lane_id();
    uint32_t s = x;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t y = __shfl_up(s, d, kWave);
        if (lane >= uint32_t(d)) s += y;
    }
    *total = __shfl(s, kWave - 1, kWave);
    return s - x;
}

// Wave-wide maximum and sum (every lane gets the result).
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long x) {
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const unsigned long long y = __shfl_xor(x, o, kWave);
        x = y > x ? y : x;
    }
    return x;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
    for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
    return x;
}

// Exclusive prefix sum over the block; s_tmp needs kWavesPerBlock + 1 words.
// Must be called by every thread of the block.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* total, uint32_t* s_tmp) {
    const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
    uint32_t wt;
    const uint32_t wex = wave_excl_scan(x, &wt);
    if (lane == 0) s_tmp[wid] = wt;
    __syncthreads();
    if (wid == 0) {
        const uint32_t v = lane < uint32_t(kWavesPerBlock) ? s_tmp[lane] : 0u;
        uint32_t t;
        const uint32_t e = wave_excl_scan(v, &t);
        if (lane < uint32_t(kWavesPerBlock)) s_tmp[lane] = e;
        if (lane == 0) s_tmp[kWavesPerBlock] = t;
    }
    __syncthreads();
    *total = s_tmp[kWavesPerBlock];
    const uint32_t r = wex + s_tmp[wid];
    __syncthreads();
    return r;
}

__device__ __forceinline__ uint64_t make_item(uint64_t start, uint32_t len) { return start | (uint64_t(len) << 40); }
__device__ __forceinline__ uint64_t item_start(uint64_t it) { return it & ((1ull << 40) - 1); }
__device__ __forceinline__ uint32_t item_len(uint64_t it) { return uint32_t(it >> 40); }

__device__ __forceinline__ uint64_t* defer_list(const Inst& in, uint32_t parity) { return parity ? in.pl1 : in.pl0; }

// Wave-aggregated append of the adjacency of the lanes' vertices (defer ==
// true) to a deferred list, in chunks of kDefer entries.  All lanes call.
__device__ __forceinline__ void defer_push(const Inst& in, uint64_t* list, unsigned long long* count, Control* ctl,
                                           bool defer, uint32_t v, unsigned long long* s_ent = nullptr) {
    uint64_t beg = 0, dt = 0;
    uint32_t nch = 0;
    if (defer) {
        beg = uint64_t(in.off[v]);
        dt = uint64_t(in.off[v + 1]) - beg;
        nch = uint32_t((dt + in.defer - 1) / in.defer);
    }
    if (__ballot(nch != 0) == 0) return;
    if (s_ent) {  // the deferred entries, for the next step's binned apply
        const unsigned long long d = wave_sum_u64(dt);
        if (lane_id() == 0) atomicAdd(s_ent, d);
    }
    uint32_t total;
    const uint32_t ex = wave_excl_scan(nch, &total);
    unsigned long long base = 0;
    if (lane_id() == 0) base = atomicAdd(count, (unsigned long long)total);
    base = __shfl(base, 0, kWave);
    if (base + total > in.plcap) {  // cannot happen: plcap bounds all deferrals
        if (lane_id() == 0) atomicCAS(&ctl->error, 0, int32_t(kErrDefer));
        return;
    }
    for (uint32_t c = 0; c < nch; ++c) {
        const uint64_t rem = dt - uint64_t(c) * in.defer;
        list[base + ex + c] = make_item(beg + uint64_t(c) * in.defer, rem < in.defer ? uint32_t(rem) : in.defer);
    }
}

// Block-aggregated append of the adjacency chunks of the threads' vertices
// (defer == true): one atomic per block.  All threads call.
__device__ __forceinline__ void block_defer_push(const Inst& in, uint64_t* list, unsigned long long* count,
                                                 Control* ctl, bool defer, uint32_t v, uint32_t* s_tmp,
                                                 unsigned long long* s_base) {
    uint64_t beg = 0, dt = 0;
    uint32_t nch = 0;
    if (defer) {
        beg = uint64_t(in.off[v]);
        dt = uint64_t(in.off[v + 1]) - beg;
        nch = uint32_t((dt + in.defer - 1) / in.defer);
    }
    uint32_t total;
    const uint32_t ex = block_excl_scan(nch, &total, s_tmp);
    if (total == 0) return;
    if (threadIdx.x == 0) *s_base = atomicAdd(count, (unsigned long long)total);
    __syncthreads();
    const unsigned long long base = *s_base;
    if (base + total > in.plcap) {  // cannot happen: plcap bounds all deferrals
        if (threadIdx.x == 0) atomicCAS(&ctl->error, 0, int32_t(kErrDefer));
        return;
    }
    for (uint32_t c = 0; c < nch; ++c) {
        const uint64_t rem = dt - uint64_t(c) * in.defer;
        list[base + ex + c] = make_item(beg + uint64_t(c) * in.defer, rem < in.defer ? uint32_t(rem) : in.defer);
    }
}

// The same append for U vertices per thread (defer[r] for v[r]): one block
// scan and one atomic for all of them.  All threads call.
template <uint32_t U>
__device__ __forceinline__ void block_defer_push_n(const Inst& in, uint64_t* list, unsigned long long* count,
                                                   Control* ctl, const bool* defer, const uint32_t* v,
                                                   uint32_t* s_tmp, unsigned long long* s_base) {
    uint32_t nch = 0;
#pragma unroll
    for (uint32_t r = 0; r < U; ++r)
        if (defer[r]) nch += uint32_t((uint64_t(in.off[v[r] + 1] - in.off[v[r]]) + in.defer - 1) / in.defer);
    uint32_t total;
    const uint32_t ex = block_excl_scan(nch, &total, s_tmp);
    if (total == 0) return;
    if (threadIdx.x == 0) *s_base = atomicAdd(count, (unsigned long long)total);
    __syncthreads();
    const unsigned long long base = *s_base;
    if (base + total > in.plcap) {  // cannot happen: plcap bounds all deferrals
        if (threadIdx.x == 0) atomicCAS(&ctl->error, 0, int32_t(kErrDefer));
        return;
    }
    uint32_t at = ex;
#pragma unroll
    for (uint32_t r = 0; r < U; ++r) {
        if (!defer[r]) continue;
        const uint64_t beg = uint64_t(in.off[v[r]]);
        const uint64_t dt = uint64_t(in.off[v[r] + 1]) - beg;
        for (uint64_t c0 = 0; c0 < dt; c0 += in.defer) {
            const uint64_t rem = dt - c0;
            list[base + at++] = make_item(beg + c0, rem < in.defer ? uint32_t(rem) : in.defer);
        }
    }
}

// ----------------------------------------------------------- graph set-up
// Set-up words read back by the host (one allocation, zeroed first).
enum Prep : int { kPrepMaxDeg = 0, kPrepBad = 1, kPrepDefer = 2, kPrepHeavy = 3, kPrepMaxNz = 4, kPrepWords = 8 };
enum PrepBad : unsigned long long { kBadRowPtr = 1, kBadCol = 2 };


// 32-bit caller offsets widened to int64.
__global__ void k_widen_offsets(const int32_t* src, uint64_t count, int64_t* dst) {
    for (uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; t < count; t += uint64_t(gridDim.x) * blockDim.x)
        dst[t] = src[t];
}

// Per vertex v (caller's ids): tot[v] = entries over all layers, nz[v] =
// layers in which v has an edge, ids[v] = v (the sort's values); the
// largest layer degree and a flag for decreasing offsets go to prep.
__global__ void k_prep_vertex(const int64_t* const* row_ptr, uint32_t n, uint32_t L, uint64_t* tot, uint8_t* nz,
                              uint32_t* ids, unsigned long long* prep) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    const uint64_t first = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
    // every lane runs the same number of rounds so the wave reductions see all lanes
    for (uint64_t base = first - threadIdx.x % kWave; base < n; base += stride) {
        const uint64_t v = base + threadIdx.x % kWave;
        unsigned long long md = 0, bad = 0;
        if (v < n) {
            uint64_t t = 0;
            uint32_t z = 0;
            for (uint32_t l = 0; l < L; ++l) {
                const int64_t d = row_ptr[l][v + 1] - row_ptr[l][v];
                if (d < 0) {
                    bad = kBadRowPtr;
                    continue;
                }
                t += uint64_t(d);
                z += d > 0;
                md = uint64_t(d) > md ? uint64_t(d) : md;
            }
            tot[v] = t;
            nz[v] = uint8_t(z);
            ids[v] = uint32_t(v);
        }
        md = wave_max_u64(md);
        bad = wave_max_u64(bad);
        if (lane_id() == 0) {
            if (md) atomicMax(&prep[kPrepMaxDeg], md);
            if (bad) atomicOr(&prep[kPrepBad], bad);
        }
    }
}

// Rank r of the degree order (tot_s descending, perm[r] = caller's id):
// inverse permutation, z in internal ids, chunk counts of the static chunk
// list; the deferred-chunk bound, the heavy head (ranks with more than
// kPullHeavy entries) and the largest z go to prep.
template <uint64_t kChunkEntries, uint64_t kDeferEntries, uint64_t kHeavyEntries>
__global__ void k_prep_rank(const uint64_t* tot_s, const uint32_t* perm, const uint8_t* nz, uint32_t n, uint32_t* inv,
                            uint8_t* zr, uint64_t* chunk_cnt, unsigned long long* prep) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    const uint64_t first = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
    for (uint64_t base = first - threadIdx.x % kWave; base < n; base += stride) {
        const uint64_t r = base + threadIdx.x % kWave;
        unsigned long long defer = 0, heavy = 0, mz = 0;
        if (r < n) {
            const uint32_t v = perm[r];
            const uint64_t t = tot_s[r];
            inv[v] = uint32_t(r);
            const uint8_t z = nz[v];
            zr[r] = z;
            mz = z;
            chunk_cnt[r] = (t + kChunkEntries - 1) / kChunkEntries;
            defer = (t + kDeferEntries - 1) / kDeferEntries;
            if (t > kHeavyEntries) heavy = r + 1;  // totals descend with r
        }
        defer = wave_sum_u64(defer);
        heavy = wave_max_u64(heavy);
        mz = wave_max_u64(mz);
        if (lane_id() == 0) {
            if (defer) atomicAdd(&prep[kPrepDefer], defer);
            if (heavy) atomicMax(&prep[kPrepHeavy], heavy);
            if (mz) atomicMax(&prep[kPrepMaxNz], mz);
        }
    }
}

// deg0[v][l] = deg_l(perm[v]) from per-layer offsets (v: internal id,
// perm[v]: caller's id).
__global__ void k_build_deg0(const int64_t* const* row_ptr, const uint32_t* perm, uint32_t n, uint32_t L,
                             uint32_t* deg0) {
    const uint64_t total = uint64_t(n) * L;
    for (uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; t < total; t += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t v = perm[t / L], l = uint32_t(t % L);
        deg0[t] = uint32_t(row_ptr[l][v + 1] - row_ptr[l][v]);
    }
}

// Merged vertex-major adjacency in internal ids, and the target's z per
// entry; one wavefront per vertex.  A neighbour id >= n sets kBadCol in
// prep[kPrepBad] (the entry is written as a self-reference of no weight:
// target 0, z 0) so the host can reject the graph.
__global__ void k_build_adj(const int64_t* const* row_ptr, const uint32_t* const* col, const uint32_t* deg0,
                            const int64_t* off, const uint32_t* perm, const uint32_t* inv, const uint8_t* z_caller,
                            uint32_t n, uint32_t L, uint32_t* adj, uint8_t* adjz, unsigned long long* prep) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
    const uint32_t lane = lane_id();
    bool bad = false;
    for (uint32_t v = wave; v < n; v += nwaves) {
        const uint32_t ov = perm[v];
        uint64_t dst = uint64_t(off[v]);
        for (uint32_t l = 0; l < L; ++l) {
            const int64_t s = row_ptr[l][ov];
            const uint32_t d = deg0[uint64_t(v) * L + l];
            const uint32_t* src = col[l] + s;
            for (uint32_t i = lane; i < d; i += kWave) {
                const uint32_t c = src[i];
                const bool ok = c < n;
                bad |= !ok;
                adj[dst + i] = ok ? inv[c] * L + l : 0u;
                adjz[dst + i] = ok ? z_caller[c] : uint8_t(0);
            }
            dst += d;
        }
    }
    if (__ballot(bad) && lane == 0 && prep) atomicOr(&prep[kPrepBad], (unsigned long long)kBadCol);
}

// Per-vertex core-index bounds: hv[v * L + j] = min(h, 255), h = the
// (j+1)-th largest over layers l of d_j(v, l), the entries of v in layer l
// whose target has z > j -- the degrees instance lmd = j + 1 starts from
// (no other neighbour is ever alive there).  A vertex of the (k, j+1) core
// has j+1 layers of degree >= k, so its index is at most h; h = 0 exactly
// for the vertices that instance removes at level 0 (or never counts).  One
// wavefront per vertex: the counts by ballot, then lane j < L ranks the L
// layer counts of instance j.  L <= 16.
__global__ void k_build_vbound(const int64_t* off, const uint32_t* deg0, const uint8_t* adjz, uint32_t n, uint32_t L,
                               uint8_t* hv) {
    const uint32_t lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
    for (uint32_t v = wave; v < n; v += nwaves) {
        uint64_t pos = uint64_t(off[v]);
        uint32_t x[16];  // lane j: d_j(v, l) per layer l
#pragma unroll
        for (uint32_t l = 0; l < 16; ++l) {
            x[l] = 0;
            if (l < L) {
                const uint32_t d = deg0[uint64_t(v) * L + l];
                uint32_t mine = 0;
                for (uint32_t t0 = 0; t0 < d; t0 += kWave) {
                    const uint32_t z = t0 + lane < d ? uint32_t(adjz[pos + t0 + lane]) : 0u;
                    for (uint32_t j = 0; j < L; ++j) {
                        const uint32_t c = uint32_t(__popcll(__ballot(z > j)));
                        if (lane == j) mine += c;
                    }
                }
                x[l] = mine;
                pos += d;
            }
        }
        if (lane < L) {
            uint32_t h = 0;
#pragma unroll
            for (uint32_t l = 0; l < 16; ++l) {
                if (l < L) {
                    uint32_t r = 0;  // rank of layer l, descending, ties by layer
#pragma unroll
                    for (uint32_t m = 0; m < 16; ++m) r += uint32_t(m < L && (x[m] > x[l] || (x[m] == x[l] && m < l)));
                    if (r == lane) h = x[l];
                }
            }
            hv[uint64_t(v) * L + lane] = uint8_t(h < 255u ? h : 255u);
        }
    }
}

// adjh[j * D + p] = hv[u * L + j] for the target u of entry p and every lmd
// j + 1: the bound streams beside the entry.
__global__ void k_build_adjh(const uint32_t* adj, uint64_t D, uint32_t L, uint32_t lmagic, const uint8_t* hv,
                             uint8_t* adjh) {
    for (uint64_t p = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; p < D; p += uint64_t(gridDim.x) * blockDim.x) {
        const uint8_t* h = hv + uint64_t(div_L(adj[p], L, lmagic)) * L;
        for (uint32_t j = 0; j < L; ++j) adjh[uint64_t(j) * D + p] = h[j];
    }
}

// out[i][w] = core index of instance i at internal id inv[w]: results in
// the caller's ids.
__global__ void k_unpermute(const Inst* insts, const uint32_t* inv, uint32_t n, uint32_t I, int32_t* out) {
    const uint64_t total = uint64_t(n) * I;
    for (uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; t < total; t += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t i = t / n, w = t % n;
        out[t] = *idx_of(insts[i], inv[w]);
    }
}

// Static chunk list (vertex, chunk#) covering all adjacency in kChunk pieces:
// used by passes that visit every entry with its source vertex.
__global__ void k_build_chunks(const int64_t* off, const uint64_t* chunk_off, uint32_t n, uint64_t* chunks) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const uint64_t b = chunk_off[v], e = chunk_off[v + 1];
        for (uint64_t c = b; c < e; ++c) chunks[c] = uint64_t(v) | ((c - b) << 32);
    }
}

// ------------------------------------------------------------ step control
// What step s does for an instance, from the previous step's slots: continue
// the level while the previous cascade deferred chunks to this step, finish
// when no vertex is alive, otherwise start the next level.  Every block of
// the step's level kernel (k_init or k_scan) decides for itself: the inputs
// are not written during the kernel, except the alive count, which only
// falls -- a block that sees it reach 0 has no alive vertex left in its
// slice, and once every vertex is removed the rest of the step changes no
// index.  Block 0 records its decision for the cascade and the next step.
__device__ __forceinline__ int32_t decide(InstState* st, uint32_t s, int32_t* k) {
    const uint32_t p = (s + 1u) & 1u;  // parity of step s - 1
    const int32_t prev = st->dphase[p];
    *k = st->dk[p];
    if (prev == kPhaseDone) return kPhaseDone;
    if (st->pdefer[s & 1u][0] > 0) return kPhaseCont;
    if (s > 0 && __hip_atomic_load(&st->alive[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return kPhaseDone;
    *k += 1;
    return kPhaseScan;
}

struct StepShared {
    int32_t phase[64];
    int32_t k[64];
};

// Every thread i < I decides instance i into sh; block 0 records the
// decisions, clears the next step's counters and reports the active count.
// All threads call.
__device__ __forceinline__ void step_decide(InstState* states, Control* ctl, uint32_t I, uint32_t s,
                                            StepShared* sh) {
    if (threadIdx.x < I) {
        int32_t k;
        const int32_t ph = decide(&states[threadIdx.x], s, &k);
        sh->phase[threadIdx.x] = ph;
        sh->k[threadIdx.x] = k;
        if (blockIdx.x == 0) {
            InstState* st = &states[threadIdx.x];
            st->dphase[s & 1u] = ph;
            st->dk[s & 1u] = k;
            st->pdefer[(s + 1u) & 1u][0] = 0;
            st->pfront[(s + 1u) & 1u][0] = 0;
            st->pfent[(s + 1u) & 1u][0] = 0;
            st->bmode[(s + 1u) & 1u] = 0;
        }
    }
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int32_t active = 0;
        for (uint32_t i = 0; i < I; ++i) active += sh->phase[i] != kPhaseDone;
        ctl->active = active;
    }
}

// ------------------------------------------------------------------ scans
struct ScanShared {
    uint32_t keep;     // alive-list write position
    uint32_t removed;  // removals by this block (for the alive counter)
    uint32_t tmp[kWavesPerBlock + 1];
    unsigned long long base;
    unsigned long long ent;  // adjacency entries of the block's removals
};

// ------------------------------------------------------------- level 0
// Level 0 by pull (L <= kPullMaxL).  With z(v) = number of layers in which v
// has an edge, every vertex with z(v) < lmd leaves instance lmd at once.
// Instead of pushing their removal to all their neighbours, each vertex v
// counts, per layer l and for every lmd <= z(v) at once,
//     deg_lmd[v][l] = #{w in N_l(v) : z(w) >= lmd}
// in one pass over its adjacency (16 register counters, no atomics).  The
// level-0 scan then removes the survivors with fewer than lmd non-empty
// layers; only those are pushed (their neighbours counted them).
constexpr uint32_t kPullMaxL = 16;
// Vertices with more entries are pulled by a wavefront (k_pull_heavy); the
// others by a thread, and their per-layer counts (<= kPullHeavy) fit a byte.
constexpr uint64_t kPullHeavy = 255;
constexpr uint32_t kStageWords = 32768;  // level-0 row staging: instances * layers * kBlock <= this

// Level-0 counts: c[j] += #{entries of a segment whose target z > j},
// j < kPullMaxL, byte-parallel over 16-byte words of the z stream (z values
// <= kPullMaxL, so byte + 127 - j >= 128 iff byte > j and no byte carries
// into the next): for each j, one add, one mask and a popcount per 4 bytes.
// sel holds 0x80 in every byte that belongs to the segment.
// NJ: counters kept (instances of the graph, rounded up to 8 or 16).
template <uint32_t NJ>
__device__ __forceinline__ void count_word(uint32_t w, uint32_t sel, uint32_t* c) {
#pragma unroll
    for (uint32_t j = 0; j < NJ; ++j) c[j] += uint32_t(__popc((w + (0x7F7F7F7Fu - j * 0x01010101u)) & sel));
}

// 0x80 in each byte t of the 4-byte word starting at byte b0 with beg <= t < end.
__device__ __forceinline__ uint32_t byte_sel(uint64_t b0, uint64_t beg, uint64_t end) {
    uint32_t sel = 0;
#pragma unroll
    for (uint32_t b = 0; b < 4; ++b)
        if (b0 + b >= beg && b0 + b < end) sel |= 0x80u << (8 * b);
    return sel;
}

// c[j] += #{t in [beg, end) : z[t] > j}, reading the 16-byte words of z
// that cover the range (z 16-byte aligned and readable up to the word that
// holds end - 1), word q0 + first, q0 + first + step, ...
template <uint32_t NJ>
__device__ __forceinline__ void pull_count_vec(const uint8_t* z, uint64_t beg, uint64_t end, uint32_t first,
                                               uint32_t step, uint32_t* c) {
    if (beg >= end) return;
    const uint4* zp = reinterpret_cast<const uint4*>(z);
    const uint64_t q1 = (end - 1) >> 4;
    for (uint64_t q = (beg >> 4) + first; q <= q1; q += step) {
        const uint4 x = zp[q];
        const uint64_t b0 = q << 4;
        if (b0 >= beg && b0 + 16 <= end) {  // whole word inside the segment
            count_word<NJ>(x.x, 0x80808080u, c);
            count_word<NJ>(x.y, 0x80808080u, c);
            count_word<NJ>(x.z, 0x80808080u, c);
            count_word<NJ>(x.w, 0x80808080u, c);
        } else {  // a word at an end of the segment: only its selected dwords
            const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                const uint32_t sel = byte_sel(b0 + 4 * k, beg, end);
                if (sel) count_word<NJ>(w[k], sel, c);
            }
        }
    }
}

// The same count for one thread's segment: short segments (the common
// case, a few entries per layer) byte by byte, long ones by words.
template <uint32_t NJ>
__device__ __forceinline__ void pull_count_seg(const uint8_t* z, uint64_t beg, uint64_t end, uint32_t* c) {
    if (end - beg < 16) {
        for (uint64_t t = beg; t < end; ++t) {
            const uint32_t zw = z[t];
#pragma unroll
            for (uint32_t j = 0; j < NJ; ++j) c[j] += zw > j;
        }
    } else {
        pull_count_vec<NJ>(z, beg, end, 0, 1, c);
    }
}

// Hubs (ids < H, the heaviest vertices after the degree sort): one
// wavefront per vertex, lanes strided over the 16-byte words of each layer
// segment.
template <uint32_t NJ>
__global__ void __launch_bounds__(kBlock) k_pull_heavy(const Inst* insts, uint32_t I, const uint32_t* deg0,
                                                       const uint8_t* z, uint32_t H, uint32_t L) {
    const uint32_t lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint32_t nwaves = gridDim.x * blockDim.x / kWave;
    const Inst& in0 = insts[0];
    for (uint32_t v = wave; v < H; v += nwaves) {
        const uint32_t zv = z[v];
        uint64_t pos = uint64_t(in0.off[v]);
        for (uint32_t l = 0; l < L; ++l) {
            const uint32_t d = deg0[uint64_t(v) * L + l];
            uint32_t c[NJ];
#pragma unroll
            for (uint32_t j = 0; j < NJ; ++j) c[j] = 0;
            pull_count_vec<NJ>(in0.adjz, pos, pos + d, lane, kWave, c);
            pos += d;
#pragma unroll
            for (uint32_t j = 0; j < NJ; ++j)
                for (int o = kWave / 2; o > 0; o >>= 1) c[j] += __shfl_xor(c[j], o, kWave);
            if (lane == 0)
                for (uint32_t i = 0; i < I && i < zv; ++i) insts[i].deg[uint64_t(v) * L + l] = c[i];
        }
    }
}

// Level 0 for every instance: degree rows (pulled, or copied from deg0 when
// L > kPullMaxL), cnt, index, liveness bit and alive list; removed vertices
// that neighbours still count go to the frontier list.  H: vertices below
// it were pulled by k_pull_heavy.
// Three blocks per CU (the grid of every level kernel, all resident at
// once): six wavefronts per SIMD, so at most 80 VGPRs.
template <uint32_t NJ>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6)))
k_init(const Inst* insts, InstState* states, Control* ctl, uint32_t I,
                                                 const uint32_t* deg0, const uint8_t* z, uint32_t n, uint32_t L,
                                                 uint32_t cap, uint32_t H, uint32_t pull) {
    __shared__ ScanShared sh[64];
    __shared__ StepShared ss;
    // Pulled rows of light vertices are staged here (a byte suffices: a light
    // vertex has at most kPullHeavy entries) and written out in runs of 64
    // rows (the rows an instance never reads left out), so every line of deg
    // is written once.  32 KB of
    // staging leaves room for three blocks per CU, the grid of every level
    // kernel (all of them resident at once).
    __shared__ uint8_t s_stage[kStageWords];   // [instance][candidate][layer]
    __shared__ uint8_t s_c1[NJ * kBlock];  // layers with a pulled count >= 1
    __shared__ uint8_t s_zc[kBlock];       // z of each candidate (0 for none)
    const bool stage = pull && I * L * uint32_t(kBlock) <= kStageWords;
    for (uint32_t i = threadIdx.x; i < I; i += blockDim.x) {
        sh[i].keep = 0;
        sh[i].removed = 0;
        sh[i].ent = 0;
    }
    step_decide(states, ctl, I, 0u, &ss);
    __syncthreads();
    const uint64_t lo = uint64_t(blockIdx.x) * cap;  // this block's list slice
    const uint32_t G = gridDim.x;
    // runs owned: c = b, b+G, ... < ceil(n / kRun); cap >= their vertex count
    const uint64_t nruns = (uint64_t(n) + kRun - 1) / kRun;
    const uint32_t myruns = blockIdx.x < nruns ? uint32_t((nruns - blockIdx.x + G - 1) / G) : 0u;
    const Inst& in0 = insts[0];
    for (uint32_t base = 0; base < myruns * kRun; base += kBlock) {
        const uint32_t m = base + threadIdx.x;  // m-th candidate of this block
        const uint64_t vv = (uint64_t(m / kRun) * G + blockIdx.x) * kRun + m % kRun;
        const bool valid = m < myruns * kRun && vv < n;
        const uint32_t v = valid ? uint32_t(vv) : 0u;
        const uint32_t zv = valid ? uint32_t(z[v]) : 0u;
        if (valid && zv > 0) {
            if (!pull) {
                // copy the row once per instance lmd <= z(v)
                for (uint32_t l = 0; l < L; ++l) {
                    const uint32_t d = deg0[uint64_t(v) * L + l];
                    for (uint32_t i = 0; i < I && i < zv; ++i) rec_of(insts[i], v)[l] = d;
                }
            } else if (v >= H) {
                uint64_t pos = uint64_t(in0.off[v]);
                uint32_t c1v[NJ];
#pragma unroll
                for (uint32_t j = 0; j < NJ; ++j) c1v[j] = 0;
                for (uint32_t l = 0; l < L; ++l) {
                    const uint32_t d = deg0[uint64_t(v) * L + l];
                    uint32_t c[NJ];
#pragma unroll
                    for (uint32_t j = 0; j < NJ; ++j) c[j] = 0;
                    pull_count_seg<NJ>(in0.adjz, pos, pos + d, c);
                    pos += d;
#pragma unroll
                    for (uint32_t j = 0; j < NJ; ++j) {
                        if (j >= I) continue;
                        c1v[j] += c[j] >= 1u;
                        if (stage) s_stage[(j * kBlock + threadIdx.x) * L + l] = uint8_t(j < zv ? c[j] : 0u);
                        else if (j < zv) insts[j].deg[uint64_t(v) * L + l] = c[j];
                    }
                }
                if (stage) {
#pragma unroll
                    for (uint32_t j = 0; j < NJ; ++j)
                        if (j < I) s_c1[j * kBlock + threadIdx.x] = uint8_t(j < zv ? c1v[j] : 0u);
                }
            }
        }
        if (stage) {
            // the wavefront's 64 candidates are one run of consecutive ids:
            // its rows are 64 * L consecutive words of each instance's deg.
            // Instance j writes only the rows of vertices with z > j: no
            // other vertex is ever alive there, and no kernel reads the row
            // of a vertex that is not alive (hub rows come from k_pull_heavy)
            s_zc[threadIdx.x] = uint8_t(valid ? zv : 0u);
            __syncthreads();
            const uint32_t w0 = threadIdx.x & ~uint32_t(kWave - 1);  // first candidate of this wavefront
            const uint64_t run0 = (uint64_t((base + w0) / kRun) * G + blockIdx.x) * kRun;
            if (base + w0 < myruns * kRun && run0 < n) {
                for (uint32_t j = 0; j < I; ++j) {
                    uint32_t* rows = insts[j].deg + run0 * L;
                    for (uint32_t t = lane_id(); t < kRun * L; t += kWave) {
                        const uint64_t vt = run0 + t / L;
                        if (vt < n && vt >= H && s_zc[w0 + t / L] > j)
                            rows[t] = s_stage[(j * kBlock + w0 + t / L) * L + t % L];
                    }
                }
            }
        }
        for (uint32_t i = 0; i < I; ++i) {
            const Inst& in = insts[i];
            if (ss.phase[i] == kPhaseDone) {  // lmd above every vertex's layer count: all zero
                if (valid) *idx_of(in, v) = 0;
                continue;
            }
            static_assert(kRun == kWave, "bitmap words are written per wavefront run");
            // layers with degree >= 1 among the rows just built
            uint32_t c1 = 0;
            if (valid && zv >= in.lmd) {
                if (pull && stage && v >= H) {
                    c1 = s_c1[i * kBlock + threadIdx.x];
                } else if (pull) {
                    const uint32_t* d = rec_of(in, v);
                    for (uint32_t l = 0; l < L; ++l) c1 += d[l] >= 1u;
                } else {
                    c1 = zv;
                }
            }
            const bool counted = valid && zv >= in.lmd;  // neighbours' rows include v
            const bool alive = counted && c1 >= in.lmd;
            if (valid) {
                *idx_of(in, v) = alive ? -1 : 0;
                *cnt_of(in, v) = int32_t(c1);  // < lmd for every vertex not alive
            }
            {  // this wavefront's lanes are the 64 ids of one aligned run
                const uint64_t am = __ballot(alive);
                const uint64_t run0 = (uint64_t((base + threadIdx.x) / kRun) * G + blockIdx.x) * kRun;
                if (lane_id() == 0 && run0 < n) {
                    in.alive[run0 >> 5] = uint32_t(am);
                    if (run0 + 32 < n) in.alive[(run0 >> 5) + 1] = uint32_t(am >> 32);
                }
            }
            // pull: only vertices that neighbours counted are pushed; push:
            // every removed vertex is
            const bool push = pull ? (counted && !alive) : (valid && !alive);
            const uint64_t rm = __ballot(valid && !alive);
            if (rm && lane_id() == 0) atomicAdd(&sh[i].removed, uint32_t(__popcll(rm)));
            const uint64_t km = __ballot(alive);
            if (km) {
                uint32_t kb = 0;
                if (lane_id() == 0) kb = atomicAdd(&sh[i].keep, uint32_t(__popcll(km)));
                kb = __shfl(kb, 0, kWave);
                if (alive)
                    in.list[lo + kb + __builtin_amdgcn_mbcnt_hi(uint32_t(km >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(km), 0u))] = v;
            }
            {  // entries of the pushed vertices (the binned apply's density test)
                const uint64_t pm = __ballot(push);
                if (pm) {
                    unsigned long long d = push ? (unsigned long long)(in.off[v + 1] - in.off[v]) : 0ull;
                    d = wave_sum_u64(d);
                    if (lane_id() == 0) atomicAdd(&sh[i].ent, d);
                }
            }
            block_defer_push(in, defer_list(in, 0u), &states[i].pfront[0][0], ctl, push, v, sh[i].tmp, &sh[i].base);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < I; i += blockDim.x) {
        const Inst& in = insts[i];
        if (ss.phase[i] == kPhaseDone) continue;
        in.blk[2 * blockIdx.x] = sh[i].keep;
        if (sh[i].removed) atomicAdd(&states[i].alive[0], 0ull - (unsigned long long)sh[i].removed);
        if (sh[i].ent) atomicAdd(&states[i].pfent[0][0], sh[i].ent);
    }
}

// #{l < L : d[l] >= thr} for a degree row written by earlier kernels; rows
// of a multiple of four words are 16-byte aligned (the rows start at v * L
// words of a 256-byte aligned array) and read four words per load.
__device__ __forceinline__ uint32_t row_count_ge(const uint32_t* d, uint32_t L, uint32_t thr) {
    uint32_t c = 0;
    if ((L & 3u) == 0) {
        const uint4* q = reinterpret_cast<const uint4*>(d);
        for (uint32_t w = 0; w < L / 4; ++w) {
            const uint4 x = q[w];
            c += uint32_t(x.x >= thr) + uint32_t(x.y >= thr) + uint32_t(x.z >= thr) + uint32_t(x.w >= thr);
        }
    } else {
        for (uint32_t l = 0; l < L; ++l) c += d[l] >= thr;
    }
    return c;
}

// Level k >= 1 scan of every instance in phase Scan: walk the block's alive
// list, drop vertices removed since (idx >= 0), remove those with fewer than
// lmd layers of degree >= k+1, keep the rest with their cnt.  Each thread
// takes kScanU list entries per round, their loads issued together (the walk
// is a chain of dependent loads: list, liveness word, row).  The list is
// compacted in place (writes never pass the round being read).
constexpr uint32_t kScanU = 4;
__global__ void __launch_bounds__(kBlock) k_scan(const Inst* insts, InstState* states, Control* ctl, uint32_t I,
                                                 uint32_t s) {
    __shared__ ScanShared sh;
    __shared__ StepShared ss;
    step_decide(states, ctl, I, s, &ss);
    for (uint32_t i = 0; i < I; ++i) {
        InstState* st = &states[i];
        if (ss.phase[i] != kPhaseScan) continue;
        const Inst& in = insts[i];
        const int32_t k = ss.k[i];
        const uint32_t thr = uint32_t(k) + 1u;
        const uint32_t L = in.L;
        const uint64_t lo = uint64_t(blockIdx.x) * in.cap;
        uint32_t* list = in.list + lo;
        const uint32_t len = in.blk[2 * blockIdx.x];
        if (threadIdx.x == 0) {
            sh.keep = 0;
            sh.removed = 0;
            sh.ent = 0;
        }
        __syncthreads();
        for (uint32_t base = 0; base < len; base += kBlock * kScanU) {
            uint32_t v[kScanU], c[kScanU];
            bool live[kScanU], rem[kScanU];
#pragma unroll
            for (uint32_t r = 0; r < kScanU; ++r) {
                const uint32_t t = base + r * kBlock + threadIdx.x;
                v[r] = t < len ? list[t] : 0u;
            }
#pragma unroll
            for (uint32_t r = 0; r < kScanU; ++r) live[r] = base + r * kBlock + threadIdx.x < len && is_alive(in, v[r]);
#pragma unroll
            for (uint32_t r = 0; r < kScanU; ++r) c[r] = live[r] ? row_count_ge(rec_of(in, v[r]), L, thr) : 0u;
            unsigned long long d = 0;
#pragma unroll
            for (uint32_t r = 0; r < kScanU; ++r) {
                rem[r] = live[r] && c[r] < in.lmd;
                if (rem[r]) {
                    *idx_of(in, v[r]) = k;
                    clear_alive(in, v[r]);
                    d += (unsigned long long)(in.off[v[r] + 1] - in.off[v[r]]);
                }
                if (live[r]) *cnt_of(in, v[r]) = int32_t(c[r]);
            }
            __syncthreads();  // every entry of this round is read before compaction writes
#pragma unroll
            for (uint32_t r = 0; r < kScanU; ++r) {
                const bool keep = live[r] && !rem[r];
                const uint64_t km = __ballot(keep);
                if (km) {
                    uint32_t kb = 0;
                    if (lane_id() == 0) kb = atomicAdd(&sh.keep, uint32_t(__popcll(km)));
                    kb = __shfl(kb, 0, kWave);
                    if (keep)
                        list[kb + __builtin_amdgcn_mbcnt_hi(uint32_t(km >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(km), 0u))] =
                            v[r];
                }
                const uint64_t rm = __ballot(rem[r]);
                if (rm && lane_id() == 0) atomicAdd(&sh.removed, uint32_t(__popcll(rm)));
            }
            d = wave_sum_u64(d);
            if (d && lane_id() == 0) atomicAdd(&sh.ent, d);
            block_defer_push_n<kScanU>(in, defer_list(in, s & 1u), &st->pfront[s & 1u][0], ctl, rem, v, sh.tmp,
                                       &sh.base);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            in.blk[2 * blockIdx.x] = sh.keep;
            if (sh.removed) atomicAdd(&st->alive[0], 0ull - (unsigned long long)sh.removed);
            if (sh.ent) atomicAdd(&st->pfent[s & 1u][0], sh.ent);
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------- cascade
// Decrements are combined per target entry in a block-private LDS hash
// table and applied with one atomicSub(c) per distinct target: a hub that
// loses many neighbours in one batch takes one global atomic instead of one
// per neighbour.  The threshold crossing stays exact: the one subtraction
// with old >= k+1 > old - c is the crossing one.
constexpr uint32_t kAggSlots = 4096;
constexpr uint32_t kAggEmpty = 0xFFFFFFFFu;  // entries are < n*L <= 2^32 - 1
constexpr uint32_t kAggProbes = 8;

struct CascShared {
    int32_t phase[64];            // per instance: this step's phase, level and chunk count
    int32_t level[64];
    unsigned long long chunks[64];
    uint32_t pre[kBlock];
    uint64_t start[kBlock];
    uint32_t key[kAggSlots];
    uint32_t cnt[kAggSlots];
    uint16_t slots[kAggSlots];  // slots holding a key, in insertion order
    uint32_t tmp[kWavesPerBlock + 1];
    uint32_t head;
    uint32_t tail;
    uint32_t removed;
    uint32_t used;  // distinct keys in the table
    uint32_t heavy;  // this step's deferral threshold
    unsigned long long dent;  // entries this block deferred for the current instance
};

// Places the vertices removed by this wavefront's lanes (rem) into the
// block's ring (slots reserved by wave in LDS, never past head + cap) or,
// when heavy or when the ring is full, into the deferred list of the next
// step.  All lanes call.
__device__ __forceinline__ void emit_removed(const Inst& in, InstState* st, Control* ctl, CascShared* sh,
                                             uint32_t* ring, uint32_t next_parity, uint32_t head, bool rem,
                                             uint32_t u) {
    const uint64_t rm = __ballot(rem);
    if (rm == 0) return;
    if (lane_id() == 0) atomicAdd(&sh->removed, uint32_t(__popcll(rm)));
    const uint32_t cap = in.cap;
    uint64_t d = 0;
    if (rem) d = uint64_t(in.off[u + 1] - in.off[u]);
    const bool light = rem && d <= sh->heavy;
    const uint64_t lm = __ballot(light);
    bool placed = false;
    if (lm) {
        const uint32_t want = uint32_t(__popcll(lm));
        uint32_t base = 0xFFFFFFFFu;
        if (lane_id() == 0) {
            uint32_t t = sh->tail;
            for (;;) {
                if (t + want - head > cap) break;  // ring full: defer instead
                const uint32_t prev = atomicCAS(&sh->tail, t, t + want);
                if (prev == t) {
                    base = t;
                    break;
                }
                t = prev;
            }
        }
        base = __shfl(base, 0, kWave);
        if (base != 0xFFFFFFFFu) {
            if (light) {
                const uint32_t r = __builtin_amdgcn_mbcnt_hi(uint32_t(lm >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(lm), 0u));
                ring[(base + r) % cap] = u;
            }
            placed = light;
        }
    }
    defer_push(in, defer_list(in, next_parity), &st->pdefer[next_parity][0], ctl, rem && !placed, u, &sh->dent);
}

// Applies U decrements deg[e[r]] -= c[r] (where valid[r]) with the cascade
// rule; rem[r] = this call removed u[r].  The U subtractions are issued
// together, then the (rare) cnt updates: two dependent round trips per batch
// instead of two per entry.  The cnt subtraction that returns lmd is unique
// over the whole decomposition, so it alone decides the removal (no CAS on
// idx): every vertex that is not alive holds cnt < lmd (the scans and level
// 0 store cnt for the vertices they remove, a cascade removal leaves
// lmd - 1), and cnt only decreases between scans.
template <uint32_t U>
__device__ __forceinline__ void apply_batch(const Inst& in, const uint32_t* e, const uint32_t* u, const uint32_t* c,
                                            const bool* valid, uint32_t thr, int32_t k, bool* rem) {
    uint32_t old[U];
#pragma unroll
    for (uint32_t r = 0; r < U; ++r) old[r] = valid[r] ? atomicSub(&in.deg[e[r]], c[r]) : 0u;
    bool cross[U];
#pragma unroll
    for (uint32_t r = 0; r < U; ++r) cross[r] = valid[r] && old[r] >= thr && old[r] - c[r] < thr;
    int32_t oc[U];
#pragma unroll
    for (uint32_t r = 0; r < U; ++r) oc[r] = cross[r] ? atomicSub(&in.cnt[u[r]], 1) : 0;
#pragma unroll
    for (uint32_t r = 0; r < U; ++r) {
        rem[r] = cross[r] && oc[r] == int32_t(in.lmd);
        if (rem[r]) {
            *idx_of(in, u[r]) = k;
            clear_alive(in, u[r]);
        }
    }
}

// Applies every combined decrement in the LDS table and empties it.  All
// threads call.
__device__ __forceinline__ void agg_flush(const Inst& in, InstState* st, Control* ctl, CascShared* sh, uint32_t* ring,
                                          uint32_t next_parity, uint32_t head, uint32_t thr, int32_t k) {
    __syncthreads();
    const uint32_t used = sh->used;
    constexpr uint32_t U = 4;
    for (uint32_t t0 = 0; t0 < used; t0 += kBlock * U) {
        uint32_t e[U], u[U], c[U];
        bool valid[U], rem[U];
#pragma unroll
        for (uint32_t r = 0; r < U; ++r) {
            const uint32_t t = t0 + r * kBlock + threadIdx.x;
            valid[r] = t < used;
            e[r] = 0;
            c[r] = 0;
            if (valid[r]) {
                const uint32_t slot = sh->slots[t];
                e[r] = sh->key[slot];
                c[r] = sh->cnt[slot];
                sh->key[slot] = kAggEmpty;
                sh->cnt[slot] = 0;
            }
            u[r] = div_L(e[r], in.L, in.lmagic);
        }
        apply_batch<U>(in, e, u, c, valid, thr, k, rem);
#pragma unroll
        for (uint32_t r = 0; r < U; ++r) emit_removed(in, st, ctl, sh, ring, next_parity, head, rem[r], u[r]);
    }
    __syncthreads();
    if (threadIdx.x == 0) sh->used = 0;
    __syncthreads();
}

// Inserts one decrement of deg[e] into the LDS table; false if the probe
// sequence is full (the caller then applies it directly).
__device__ __forceinline__ bool agg_insert(CascShared* sh, uint32_t e) {
    uint32_t h = (e * 2654435761u) >> 20;  // 12-bit hash
    for (uint32_t p = 0; p < kAggProbes; ++p, h = (h + 1) & (kAggSlots - 1)) {
        uint32_t kk = sh->key[h];
        if (kk == kAggEmpty) {
            kk = atomicCAS(&sh->key[h], kAggEmpty, e);
            if (kk == kAggEmpty) {
                sh->slots[atomicAdd(&sh->used, 1u)] = uint16_t(h);
                kk = e;
            }
        }
        if (kk == e) {
            atomicAdd(&sh->cnt[h], 1u);
            return true;
        }
    }
    return false;
}

// Expands the adjacency ranges (start, len) held by the block's threads
// (len = 0 for none) and applies the cascade rule to every entry whose
// target is alive.  Each thread walks kUnroll entries per round with their
// loads issued together (the walk is latency-bound).  All threads call; the
// LDS table must be empty on entry and is empty on return.
constexpr uint32_t kUnroll = 4;
__device__ __forceinline__ void cascade_expand(const Inst& in, InstState* st, Control* ctl, CascShared* sh,
                                               uint32_t* ring, uint32_t next_parity, int32_t k, uint64_t start,
                                               uint32_t len) {
    uint32_t total;
    const uint32_t pre = block_excl_scan(len, &total, sh->tmp);
    sh->pre[threadIdx.x] = pre;
    sh->start[threadIdx.x] = start;
    __syncthreads();
    const uint32_t L = in.L, magic = in.lmagic;
    const uint32_t thr = uint32_t(k) + 1u;
    const uint32_t head = sh->head;
    const bool agg = (in.flags & kFlagAggregate) != 0;
    const bool compact = (in.flags & kFlagCompact) != 0;
    const EntryFilter ef = entry_filter(in, k);
    for (uint32_t j0 = 0; j0 < total; j0 += kBlock * kUnroll) {
        uint32_t e[kUnroll], u[kUnroll];
        bool live[kUnroll];
#pragma unroll
        for (uint32_t r = 0; r < kUnroll; ++r) {
            const uint32_t j = j0 + r * kBlock + threadIdx.x;
            live[r] = j < total;
            e[r] = 0;
            if (live[r]) {
                uint32_t lo = 0, hi = kBlock - 1;  // largest p with pre[p] <= j
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (sh->pre[mid] <= j) lo = mid;
                    else hi = mid - 1;
                }
                const uint64_t p = sh->start[lo] + (j - sh->pre[lo]);
                e[r] = in.adj[p];
                live[r] = compact || ef.z[p] >= ef.t;
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < kUnroll; ++r) {
            u[r] = div_L(e[r], L, magic);
            if (live[r]) live[r] = is_alive(in, u[r]);
        }
        bool rem[kUnroll], direct[kUnroll];
        uint32_t one[kUnroll];
#pragma unroll
        for (uint32_t r = 0; r < kUnroll; ++r) {
            one[r] = 1u;
            direct[r] = live[r] && !(agg && agg_insert(sh, e[r]));
        }
        apply_batch<kUnroll>(in, e, u, one, direct, thr, k, rem);
#pragma unroll
        for (uint32_t r = 0; r < kUnroll; ++r) emit_removed(in, st, ctl, sh, ring, next_parity, head, rem[r], u[r]);
        if (agg) agg_flush(in, st, ctl, sh, ring, next_parity, head, thr, k);
    }
    sync_stores();  // ring entries written above are read by other waves next
}

// Cascade of every instance not done: the chunks of this step's lists, then
// the block's ring until it is empty.  The chunks of all instances form one
// sequence (instance after instance) dealt round robin over the blocks, so a
// step with fewer chunks than blocks gives each block at most one chunk in
// total instead of one per instance: the tail of a decomposition is a long
// series of such steps, and each chunk is a chain of dependent memory round
// trips.  A block without chunks of an instance has nothing in its ring for
// it either.
__global__ void __launch_bounds__(kBlock) k_cascade(const Inst* insts, InstState* states, Control* ctl, uint32_t I,
                                                    uint32_t s) {
    __shared__ CascShared sh;
    const uint32_t G = gridDim.x, b = blockIdx.x;
    const uint32_t par = s & 1u, npar = par ^ 1u;
    for (uint32_t t = threadIdx.x; t < kAggSlots; t += kBlock) {
        sh.key[t] = kAggEmpty;
        sh.cnt[t] = 0;
    }
    if (threadIdx.x == 0) sh.used = 0;
    // the step's slots of every instance in one parallel load (written by
    // the preceding level kernel, read-only here)
    if (threadIdx.x < I) {
        const InstState* st = &states[threadIdx.x];
        sh.phase[threadIdx.x] = st->dphase[par];
        sh.level[threadIdx.x] = st->dk[par];
        // one of pdefer / pfront is zero; after a binned apply the list holds
        // the vertices it removed instead of the frontier
        sh.chunks[threadIdx.x] =
            st->bmode[par] ? st->pbr[st->bround[par] & 1][0] : st->pdefer[par][0] + st->pfront[par][0];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // A light step (few chunks in all, most blocks idle: the tail) keeps
        // heavier removals in the discovering block's ring rather than paying
        // a further step for them.
        unsigned long long tot = 0;
        for (uint32_t i = 0; i < I; ++i)
            if (sh.phase[i] != kPhaseDone) tot += sh.chunks[i];
        sh.heavy = tot <= insts[0].tail_chunks ? insts[0].tail_heavy : insts[0].heavy;
    }
    __syncthreads();
    uint64_t seq = 0;  // chunks of the instances before i
    for (uint32_t i = 0; i < I; ++i) {
        InstState* st = &states[i];
        if (sh.phase[i] == kPhaseDone) continue;
        const uint64_t pc = sh.chunks[i];
        if (pc == 0) continue;  // empty frontier: the ring starts empty, nothing to do
        // chunk j of instance i is item seq + j of the sequence; it goes to
        // block (seq + j) mod G, so this block's first chunk is j = first
        const uint64_t first = (uint64_t(b) + G - seq % G) % G;
        seq += pc;
        if (first >= pc) continue;
        const Inst& in = insts[i];
        const int32_t k = sh.level[i];
        uint32_t* ring = in.ring + uint64_t(b) * in.cap;
        if (threadIdx.x == 0) {
            sh.head = 0;
            sh.tail = 0;
            sh.removed = 0;
            sh.dent = 0;
        }
        __syncthreads();
        // frontier / deferred chunks
        const uint64_t* pl = defer_list(in, par);
        for (uint64_t m0 = 0; m0 * G + first < pc; m0 += kBlock) {
            const uint64_t j = (m0 + threadIdx.x) * G + first;
            uint64_t start = 0;
            uint32_t len = 0;
            if (j < pc) {
                const uint64_t it = pl[j];
                start = item_start(it);
                len = item_len(it);
            }
            cascade_expand(in, st, ctl, &sh, ring, npar, k, start, len);
        }
        // ring
        for (;;) {
            const uint32_t head = sh.head, tail = sh.tail;
            if (head >= tail) break;
            const uint32_t nb = tail - head < uint32_t(kBlock) ? tail - head : uint32_t(kBlock);
            uint64_t start = 0;
            uint32_t len = 0;
            if (threadIdx.x < nb) {
                const uint32_t v = ld_agent(&ring[(head + threadIdx.x) % in.cap]);
                start = uint64_t(in.off[v]);
                len = uint32_t(uint64_t(in.off[v + 1]) - start);
            }
            cascade_expand(in, st, ctl, &sh, ring, npar, k, start, len);
            if (threadIdx.x == 0) sh.head = head + nb;
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (sh.removed) atomicAdd(&st->alive[0], 0ull - (unsigned long long)sh.removed);
            if (sh.dent) atomicAdd(&st->pfent[npar][0], sh.dent);
        }
        __syncthreads();
    }
}

// Histogram of core indices per instance: hist[i*kcap + k] = #{v : idx_i(v) == k}.
constexpr int kHistLds = 4096;
__global__ void __launch_bounds__(kBlock) k_idx_hist(const int32_t* idx, uint32_t n, uint32_t I, uint64_t* hist,
                                                     uint32_t kcap) {
    __shared__ uint32_t s_h[kHistLds];
    const uint32_t i = blockIdx.y;
    const uint32_t lds_k = kcap < uint32_t(kHistLds) ? kcap : uint32_t(kHistLds);
    for (uint32_t t = threadIdx.x; t < lds_k; t += blockDim.x) s_h[t] = 0;
    __syncthreads();
    const int32_t* row = idx + uint64_t(i) * n;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const int32_t x = row[v];
        // every vertex has 0 <= x < kcap after a decomposition; anything else
        // is left out, so the host's total check reports it
        const uint32_t k = x < 0 ? kcap : uint32_t(x);
        if (k < lds_k) atomicAdd(&s_h[k], 1u);
        else if (k < kcap) atomicAdd(reinterpret_cast<unsigned long long*>(&hist[uint64_t(i) * kcap + k]), 1ull);
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < lds_k; t += blockDim.x)
        if (s_h[t]) atomicAdd(reinterpret_cast<unsigned long long*>(&hist[uint64_t(i) * kcap + t]), (unsigned long long)s_h[t]);
}

// ------------------------------------------------------- row compaction
// Between steps, when an instance has no deferred chunks pending (a level
// boundary), its rows can be rebuilt with only the entries whose target is
// still alive.  Late in a decomposition the surviving vertices are hubs
// whose rows are mostly dead neighbours (on the 100M-vertex workload about
// 90 % of the entries the tail visits), and every removal walks its whole
// row.  At a level boundary deg[v][l] of an alive v is exactly its number
// of alive neighbours in layer l (each removed vertex decremented it once),
// so the new row length is known in advance: off' = exclusive prefix sum of
// sum_l deg[v][l] over alive v.  The rows are copied in chunks of
// kCmpChunk entries, one wavefront per chunk, each reserving its places in
// the vertex's new row with one atomic on a per-vertex cursor: the order of
// a row's entries changes, which no kernel depends on.  The copy is then
// checked against the degrees (every cursor must end at off'[v+1]).
constexpr uint32_t kCmpChunk = 256;

// newlen[v] = sum_l deg[v][l] and nch[v] = chunks of its current row for
// alive v, 0 otherwise; index n gets 0 (the prefix sums run over n + 1
// items); only the sum when !kWrite.  words[0] += the sum of newlen (one
// atomic per block).
template <bool kWrite>
__global__ void __launch_bounds__(kBlock) k_cmp_count(const Inst in, uint64_t* newlen, uint64_t* nch,
                                                      unsigned long long* words) {
    __shared__ unsigned long long s_sum[kWavesPerBlock];
    unsigned long long a = 0;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t v = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; v <= in.n; v += stride) {
        uint64_t len = 0, ch = 0;
        if (v < in.n && is_alive(in, uint32_t(v))) {
            const uint32_t* d = rec_of(in, uint32_t(v));
            for (uint32_t l = 0; l < in.L; ++l) len += d[l];
            ch = (uint64_t(in.off[v + 1] - in.off[v]) + kCmpChunk - 1) / kCmpChunk;
        }
        if (kWrite) {
            newlen[v] = len;
            nch[v] = ch;
        }
        a += len;
    }
    a = wave_sum_u64(a);
    if (lane_id() == 0) s_sum[threadIdx.x / kWave] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kWavesPerBlock; ++w) t += s_sum[w];
        if (t) atomicAdd(&words[0], t);
    }
}

// Chunk list (v, c) of the alive rows, and each alive vertex's cursor set to
// the start of its new row.
__global__ void k_cmp_items(const Inst in, const int64_t* new_off, const uint64_t* item_off, uint64_t* items,
                            uint64_t* cursor) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t v = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; v < in.n; v += stride) {
        const uint64_t b = item_off[v], e = item_off[v + 1];
        for (uint64_t c = b; c < e; ++c) items[c] = v | ((c - b) << 32);
        cursor[v] = uint64_t(new_off[v]);
    }
}

// One wavefront per chunk: the entries whose target is alive go to the
// vertex's new row at places reserved with one atomic on its cursor; a
// reservation past the row's end (cannot happen when the degrees are
// exact) is not written and flags the error.
__global__ void __launch_bounds__(kBlock) k_cmp_fill(const Inst in, const uint64_t* items, uint64_t nitems,
                                                     const int64_t* new_off, uint64_t* cursor, uint32_t* new_adj,
                                                     Control* ctl) {
    const uint32_t lane = lane_id();
    const uint64_t wave = (blockIdx.x * uint64_t(blockDim.x) + threadIdx.x) / kWave;
    const uint64_t nwaves = uint64_t(gridDim.x) * blockDim.x / kWave;
    constexpr uint32_t P = kCmpChunk / kWave;
    for (uint64_t it = wave; it < nitems; it += nwaves) {
        const uint64_t x = items[it];
        const uint32_t v = uint32_t(x);
        const uint64_t c = x >> 32;
        const uint64_t beg = uint64_t(in.off[v]) + c * kCmpChunk, end = uint64_t(in.off[v + 1]);
        uint32_t e[P];
        bool keep[P];
#pragma unroll
        for (uint32_t q = 0; q < P; ++q) {
            const uint64_t p = beg + q * kWave + lane;
            keep[q] = p < end;
            e[q] = keep[q] ? in.adj[p] : 0u;
        }
        uint32_t cntq[P], total = 0;
        uint64_t mask[P];
#pragma unroll
        for (uint32_t q = 0; q < P; ++q) {
            if (keep[q]) keep[q] = is_alive(in, div_L(e[q], in.L, in.lmagic));
            mask[q] = __ballot(keep[q]);
            cntq[q] = uint32_t(__popcll(mask[q]));
            total += cntq[q];
        }
        if (total == 0) continue;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(reinterpret_cast<unsigned long long*>(&cursor[v]), (unsigned long long)total);
        base = __shfl(base, 0, kWave);
        const uint64_t row_end = uint64_t(new_off[v + 1]);
        if (base + total > row_end) {
            if (lane == 0) atomicCAS(&ctl->error, 0, int32_t(kErrCompact));
            continue;
        }
        uint64_t at = base;
#pragma unroll
        for (uint32_t q = 0; q < P; ++q) {
            if (keep[q]) {
                const uint32_t r = __builtin_amdgcn_mbcnt_hi(uint32_t(mask[q] >> 32),
                                                             __builtin_amdgcn_mbcnt_lo(uint32_t(mask[q]), 0u));
                new_adj[at + r] = e[q];
            }
            at += cntq[q];
        }
    }
}

// Every alive vertex's cursor must have reached the end of its new row.
__global__ void k_cmp_check(const Inst in, const int64_t* new_off, const uint64_t* cursor, Control* ctl) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t v = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; v < in.n; v += stride)
        if (is_alive(in, uint32_t(v)) && cursor[v] != uint64_t(new_off[v + 1])) atomicCAS(&ctl->error, 0, int32_t(kErrCompact));
}

// ------------------------------------------------------- densest subgraph
// Edge histogram of one instance: for every directed entry v -> (u, l) with
// m = min(idx[v], idx[u]) >= 1, hist[l][m] += 1.  Then num_edge_l(k) =
// sum_{m >= k} hist[l][m] = sum over v in S(k) of |N_l(v) ∩ S(k)|.
// Block-private LDS bins for small m, global atomics above.
constexpr int kLdsBins = 8192;
constexpr int kHistBlock = 256;
constexpr int kHistWaves = kHistBlock / kWave;

__global__ void __launch_bounds__(kHistBlock) k_edge_hist(const Inst in, const uint64_t* chunks, uint64_t nchunks,
                                                          uint64_t* hist, uint32_t kcap) {
    __shared__ uint32_t s_hist[kLdsBins];
    __shared__ uint32_t s_pre_all[kHistWaves][kWave];
    __shared__ uint64_t s_start_all[kHistWaves][kWave];
    __shared__ uint32_t s_src_all[kHistWaves][kWave];
    const uint32_t L = in.L;
    const uint32_t lds_k = kLdsBins / L < kcap ? kLdsBins / L : kcap;  // bins m < lds_k live in LDS
    for (uint32_t t = threadIdx.x; t < kLdsBins; t += blockDim.x) s_hist[t] = 0;
    __syncthreads();
    const uint32_t lane = lane_id();
    const uint32_t wib = threadIdx.x / kWave;
    uint32_t* s_pre = s_pre_all[wib];
    uint64_t* s_start = s_start_all[wib];
    uint32_t* s_src = s_src_all[wib];
    const uint64_t gwave = blockIdx.x * uint64_t(kHistWaves) + wib;
    const uint64_t nwaves = uint64_t(gridDim.x) * kHistWaves;
    for (uint64_t c0 = gwave * kWave; c0 < nchunks; c0 += nwaves * kWave) {
        const uint64_t c = c0 + lane;
        uint32_t len = 0, v = 0;
        uint64_t s = 0;
        if (c < nchunks) {
            const uint64_t ch = chunks[c];
            v = uint32_t(ch);
            const uint64_t ci = ch >> 32;
            const uint64_t b = uint64_t(in.off[v]), e = uint64_t(in.off[v + 1]);
            s = b + ci * kChunk;
            len = uint32_t((e - s) < uint64_t(kChunk) ? (e - s) : uint64_t(kChunk));
        }
        const int32_t iv = len ? *idx_of(in, v) : 0;
        if (iv < 1) len = 0;  // v outside S(1): contributes to no cell
        uint32_t total;
        const uint32_t pre = wave_excl_scan(len, &total);
        s_pre[lane] = pre;
        s_start[lane] = s;
        s_src[lane] = uint32_t(iv);
        __builtin_amdgcn_wave_barrier();
        for (uint32_t j0 = 0; j0 < total; j0 += kWave) {
            const uint32_t j = j0 + lane;
            if (j < total) {
                uint32_t lo = 0, hi = kWave - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (s_pre[mid] <= j) lo = mid;
                    else hi = mid - 1;
                }
                const uint32_t e = in.adj[s_start[lo] + (j - s_pre[lo])];
                const uint32_t u = div_L(e, L, in.lmagic);
                const uint32_t l = e - u * L;
                const int32_t iu = *idx_of(in, u);
                const uint32_t m = uint32_t(iu < int32_t(s_src[lo]) ? iu : int32_t(s_src[lo]));
                if (iu >= 1) {
                    if (m < lds_k) atomicAdd(&s_hist[l * lds_k + m], 1u);
                    else atomicAdd(reinterpret_cast<unsigned long long*>(&hist[uint64_t(l) * kcap + m]), 1ull);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < lds_k * L; t += blockDim.x) {
        const uint32_t x = s_hist[t];
        if (x) {
            const uint32_t l = t / lds_k, m = t % lds_k;
            atomicAdd(reinterpret_cast<unsigned long long*>(&hist[uint64_t(l) * kcap + m]), (unsigned long long)x);
        }
    }
}

// One thread per cell (k, lmd), 1 <= k <= kmax: FCPathLevelLeft::constructCore's
// wds block (Core/FCPathLevelLeft.cpp:57-99) on suffix sums of the histogram.
// hist rows are [l][m] per instance; cell_len[k] = |S(k)|.
__global__ void k_cell_score(const uint64_t* hist, uint32_t kcap, uint32_t L, uint32_t kmax, const int64_t* cell_len,
                             const double* pow_tab, double* density, int32_t* sel) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (k > kmax) return;
    double dl[64];
    for (uint32_t l = 0; l < L; ++l) {
        uint64_t s = 0;
        for (uint32_t m = k; m <= kmax; ++m) s += hist[uint64_t(l) * kcap + m];
        const uint32_t num_edge = uint32_t(s);  // uint32 accumulation as the reference
        dl[l] = double(float(num_edge) * 1.0f / 2);
    }
    // insertion sort, descending
    for (uint32_t a = 1; a < L; ++a) {
        const double x = dl[a];
        uint32_t b = a;
        while (b > 0 && dl[b - 1] < x) {
            dl[b] = dl[b - 1];
            --b;
        }
        dl[b] = x;
    }
    const double length = double(uint32_t(cell_len[k]));
    double best = 0.0;
    int32_t bsel = 0;
    for (uint32_t l = 0; l < L; ++l) {
        const double mad = (dl[l] * 1.0f) / length * 1.0;
        const double dens = mad * pow_tab[l];
        if (dens >= best) {
            best = dens;
            bsel = int32_t(l + 1);
        }
    }
    density[k] = best;
    sel[k] = bsel;
}

}  // namespace
}  // namespace hg
