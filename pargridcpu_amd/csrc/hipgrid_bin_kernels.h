// hipgrid_bin_kernels.h -- binned apply of a level's dense removal rounds.
//
// The first rounds of a level (the vertices its scan removes, F0, and those
// F0 removes, F1, ...) carry most of a decomposition's decrements: on the
// 100M-vertex workload a level's F0 decrements 10-35 % of the alive
// vertices' degree words at once.  Applied
// as scattered device-scope atomics (one 128-byte line request each, ~18 G/s
// on a 3.2 GB array whether or not the old value is returned) they are bound
// by the atomic request rate.  Here they become records (the target word
// u*L + l), moved by streaming passes, and each target range is applied by
// one workgroup in LDS:
//   k_bin_expand   decides which instances bin this step (frontier entries
//                  against alive words); expands the F0 chunks once, writes
//                  every entry whose target is alive as a record (expansion
//                  order) and counts the records per coarse bin (a coarse bin
//                  is a range of vertices holding about D / C adjacency
//                  entries, so the hub ranges are short);
//   k_bin_scatter  streams the records into their coarse bins' regions (each
//                  batch counting-sorted by bin in LDS, one global
//                  reservation per bin and batch, coalesced runs);
//   k_bin_group    one workgroup per (instance, coarse bin): groups the
//                  region's records by fine bin and publishes each fine
//                  bin's record range;
//   k_bin_fine     the fine bins (vb vertices whose degree words fit in
//                  LDS) dealt interleaved over the grid: a workgroup counts
//                  a fine bin's records per degree word with LDS atomics,
//                  applies each word's count to the row it owns and updates
//                  cnt from the layers that fell below the level's
//                  threshold: a vertex left with fewer than lmd layers at or
//                  above it is removed (index k).  Its work is proportional
//                  to the records, not to the vertices of the range.
// The vertices removed by the apply (F1) go to the step's chunk list and the
// cascade expands them as it would the frontier; from there the level runs
// as before.  The result is the same: each degree word loses exactly one
// per removed neighbour, and a vertex leaves at level k exactly when its
// count of layers at or above k+1 falls below lmd, whichever round it falls
// in.  cnt stays exact for every alive vertex (it loses exactly the layers
// that crossed), and every removed vertex holds cnt < lmd, as the cascade's
// crossing rule needs.
#pragma once

#include "hipgrid_kernels.h"

namespace hg {
namespace {

constexpr uint32_t kBinMaxC = 1024;         // coarse bins (LDS tables of the count and scatter kernels)
constexpr uint32_t kBinMaxF = kBlock;       // fine bins per coarse bin (one per thread in the prefix sum)
// LDS sizes by blocks per CU: the scatter's records per thread and batch
// (its sort buffer) and the apply's decrement counters of one fine bin; 2
// blocks per CU take the large ones.
constexpr uint32_t kBinScatterU2 = 22, kBinScatterU3 = 10;
constexpr uint32_t kBinDeltaWords2 = 16384, kBinDeltaWords3 = 9216, kBinDeltaWords4 = 6144;
constexpr uint32_t kBinUnroll = 4;
constexpr uint32_t kBinFineU = 8;  // the apply's records per thread and batch when grouping by fine bin

struct BinGeo {
    const uint32_t* cbw;  // [C+1] coarse bin boundaries as word indices (vertex * L): 0 = cbw[0] < ... < cbw[C] = n*L
    uint32_t* bh;         // [I][cmax] F0 entries per coarse bin (an upper bound of its records); zero between uses
    uint32_t* bc;         // [I][cmax] records written per coarse bin; zero between uses
    uint64_t* bb;         // [I][cmax] first record of each coarse bin's region
    uint32_t* rec1;       // [rcap] records grouped by coarse bin
    uint32_t* rec2;       // [rcap] records grouped by fine bin within each coarse bin's region
    uint64_t rcap;
    uint64_t min_ent;     // an instance bins a step only with at least this many frontier entries
    const uint16_t* gsel; // [ng+1] lowest coarse bin overlapping word granule g (words g << gsh ...); gsel[ng] = C-1
    uint32_t ng, gsh;
    uint32_t C, cmax;
    uint32_t vbw;         // degree words per fine bin (vertices per fine bin * L)
    uint32_t dens_num, dens_den;  // ... and frontier entries * den >= num * alive vertices * L
    uint32_t cont;                // continuation steps (deferred hub chunks) may bin too
    uint32_t claim_q;             // a fine bin of nv vertices is applied record by record below nv * claim_q / 4 records
    const uint32_t* ffirst;       // [C+1] first fine bin (index over all coarse bins) of each coarse bin
    const uint16_t* fcoarse;      // [nfine] coarse bin of each fine bin
    uint2* fse;                   // [I][nfine] each fine bin's records [x, y) within its coarse bin's region
    uint32_t nfine;
};

struct BinDecide {
    int32_t mode[64];
    unsigned long long rbase[64];
};

// Removal round r's input: the scan's frontier (r = 0) or the vertices round
// r - 1 removed, which exists when round r - 1 binned.
__device__ __forceinline__ bool bin_input(const InstState* st, uint32_t par, uint32_t r, bool cont,
                                          unsigned long long* ent, unsigned long long* chunks) {
    const int32_t ph = st->dphase[par];
    if (r == 0) {  // a scan step's frontier, or the hub chunks deferred to a continuation step
        *ent = st->pfent[par][0];
        *chunks = ph == kPhaseScan ? st->pfront[par][0] : st->pdefer[par][0];
        return ph == kPhaseScan || (ph == kPhaseCont && cont);
    }
    *ent = st->pber[(r - 1) & 1][0];
    *chunks = st->pbr[(r - 1) & 1][0];
    return (ph == kPhaseScan || ph == kPhaseCont) && st->bflag[(r - 1) & 1] != 0;
}

// Which instances bin round r of this step, and their record slices
// (instance order, while the buffers hold their input's entries).  Every
// block computes the same answer from words no kernel writes meanwhile.  All
// threads call.
__device__ __forceinline__ void bin_decide(const InstState* states, uint32_t I, uint32_t par, uint32_t r,
                                           const BinGeo& bg, uint32_t L, BinDecide* d) {
    if (threadIdx.x == 0) {
        unsigned long long base = 0;
        for (uint32_t i = 0; i < I; ++i) {
            const InstState* st = &states[i];
            int32_t mode = 0;
            unsigned long long rb = 0, ent = 0, chunks = 0;
            if (bin_input(st, par, r, bg.cont != 0, &ent, &chunks)) {
                const unsigned long long al = st->alive[0];
                if (ent >= bg.min_ent && ent * bg.dens_den >= (unsigned long long)bg.dens_num * al * L &&
                    base + ent <= bg.rcap) {
                    mode = 1;
                    rb = base;
                    base += ent;
                }
            }
            d->mode[i] = mode;
            d->rbase[i] = rb;
        }
    }
    __syncthreads();
}

constexpr uint32_t kBinMaxG = 4096;  // word granules of the coarse-bin lookup table

// Coarse bin of target word e: the largest c < C with cb[c] <= e, searched
// among the bins that overlap e's granule (one bin in the sparse ranges).
__device__ __forceinline__ uint32_t coarse_of(const uint32_t* cb, const uint16_t* gs, uint32_t gsh, uint32_t e) {
    const uint32_t g = e >> gsh;
    uint32_t lo = gs[g], hi = gs[g + 1];
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (cb[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// A block's batch of adjacency ranges (one per thread, length 0 for none):
// exclusive prefix of the lengths and the starts.
struct BatchShared {
    uint32_t pre[kBlock];
    uint64_t start[kBlock];
    uint32_t tmp[kWavesPerBlock + 1];
};

// Fills the batch tables; returns the batch's entry count.  All threads call.
__device__ __forceinline__ uint32_t batch_load(BatchShared* bs, uint64_t start, uint32_t len) {
    uint32_t total;
    const uint32_t pre = block_excl_scan(len, &total, bs->tmp);
    bs->pre[threadIdx.x] = pre;
    bs->start[threadIdx.x] = start;
    __syncthreads();
    return total;
}

// Adjacency position of entry j < total of the batch.
__device__ __forceinline__ uint64_t batch_pos(const BatchShared* bs, uint32_t j) {
    uint32_t lo = 0, hi = kBlock - 1;  // largest p with pre[p] <= j (its range is non-empty)
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (bs->pre[mid] <= j) lo = mid;
        else hi = mid - 1;
    }
    return bs->start[lo] + (j - bs->pre[lo]);
}

// The frontier chunk this thread takes in round m0 of the chunks dealt
// round robin over the grid (as the cascade deals them).
__device__ __forceinline__ void frontier_chunk(const uint64_t* pl, uint64_t pc, uint64_t m0, uint64_t* start,
                                               uint32_t* len) {
    const uint64_t j = (m0 + threadIdx.x) * gridDim.x + blockIdx.x;
    *start = 0;
    *len = 0;
    if (j < pc) {
        const uint64_t it = pl[j];
        *start = item_start(it);
        *len = item_len(it);
    }
}

// Expand: EU entries per thread in flight per round; records gathered in LDS
// (OUTCAP = twice a round's) before one reservation.
template <uint32_t EU>
struct BinExpandShared {
    static constexpr uint32_t kOutCap = 2 * kBlock * EU;
    BinDecide d;
    uint32_t cb[kBinMaxC + 1];
    uint16_t gs[kBinMaxG + 1];
    uint32_t hist[kBinMaxC];
    BatchShared bs;
    uint32_t out[kOutCap];
    uint32_t nout;
    unsigned long long base;
};

// Writes the gathered records to the instance's slice (one reservation) and
// empties the buffer.  All threads call.
template <uint32_t EU>
__device__ __forceinline__ void bin_out_flush(BinExpandShared<EU>* sh, InstState* st, Control* ctl, uint32_t rp,
                                              uint32_t* rec0, unsigned long long lim) {
    const uint32_t n = sh->nout;
    if (threadIdx.x == 0 && n) sh->base = atomicAdd(&st->r0cnt[rp][0], (unsigned long long)n);
    __syncthreads();
    if (n) {
        const unsigned long long b = sh->base;
        if (b + n > lim) {  // cannot happen: records are a subset of the frontier's entries
            if (threadIdx.x == 0) atomicCAS(&ctl->error, 0, int32_t(kErrBin));
        } else {
            for (uint32_t t = threadIdx.x; t < n; t += kBlock) rec0[b + t] = sh->out[t];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) sh->nout = 0;
    __syncthreads();
}

// Decision, then per binned instance: the F0 entries whose target is alive
// become records, written in expansion order to the instance's slice of the
// second record buffer (one reservation per block and batch), and bh[i][c]
// counts them per coarse bin.
template <uint32_t EU>
__global__ void __launch_bounds__(kBlock) k_bin_expand(const Inst* insts, InstState* states, Control* ctl,
                                                       uint32_t I, uint32_t s, uint32_t round, BinGeo bg) {
    __shared__ BinExpandShared<EU> sh;
    const uint32_t par = s & 1u, rp = round & 1u;
    for (uint32_t t = threadIdx.x; t <= bg.C; t += kBlock) sh.cb[t] = bg.cbw[t];
    for (uint32_t t = threadIdx.x; t <= bg.ng; t += kBlock) sh.gs[t] = bg.gsel[t];
    bin_decide(states, I, par, round, bg, insts[0].L, &sh.d);
    if (blockIdx.x == 0 && threadIdx.x < I) {
        // the round's flag, and the counters its apply fills (read by the
        // next round or the cascade; nothing reads them during this kernel)
        InstState* st = &states[threadIdx.x];
        st->bflag[rp] = sh.d.mode[threadIdx.x];
        if (sh.d.mode[threadIdx.x]) {
            st->bmode[par] = 1;
            st->bround[par] = int32_t(round);
            st->rbase[par] = sh.d.rbase[threadIdx.x];
            st->pbr[rp][0] = 0;
            st->pber[rp][0] = 0;
            atomicAdd(&ctl->binned, 1);
        }
    }
    for (uint32_t i = 0; i < I; ++i) {
        if (!sh.d.mode[i]) continue;
        const Inst& in = insts[i];
        InstState* st = &states[i];
        const bool compact = (in.flags & kFlagCompact) != 0;
        const EntryFilter ef = entry_filter(in, st->dk[par]);
        const uint64_t rb = sh.d.rbase[i];
        unsigned long long lim, pc;  // the slice holds lim records: the input's entries
        bin_input(st, par, round, bg.cont != 0, &lim, &pc);
        uint32_t* rec0 = bg.rec2 + rb;
        for (uint32_t t = threadIdx.x; t < bg.C; t += kBlock) sh.hist[t] = 0;
        if (threadIdx.x == 0) sh.nout = 0;
        __syncthreads();
        const uint64_t* pl = defer_list(in, par);
        for (uint64_t m0 = 0; m0 * gridDim.x + blockIdx.x < pc; m0 += kBlock) {
            uint64_t start;
            uint32_t len;
            frontier_chunk(pl, pc, m0, &start, &len);
            const uint32_t total = batch_load(&sh.bs, start, len);
            for (uint32_t j0 = 0; j0 < total; j0 += kBlock * EU) {
                uint32_t e[EU];
                bool ok[EU];
#pragma unroll
                for (uint32_t r = 0; r < EU; ++r) {
                    const uint32_t j = j0 + r * kBlock + threadIdx.x;
                    ok[r] = j < total;
                    e[r] = 0;
                    if (ok[r]) {
                        const uint64_t p = batch_pos(&sh.bs, j);
                        e[r] = in.adj[p];
                        ok[r] = compact || ef.z[p] >= ef.t;
                    }
                }
#pragma unroll
                for (uint32_t r = 0; r < EU; ++r) {
                    if (ok[r]) ok[r] = is_alive(in, div_L(e[r], in.L, in.lmagic));
                    if (ok[r]) atomicAdd(&sh.hist[coarse_of(sh.cb, sh.gs, bg.gsh, e[r])], 1u);
                }
                // gather in LDS (wave-aggregated slots)
#pragma unroll
                for (uint32_t r = 0; r < EU; ++r) {
                    const uint64_t m = __ballot(ok[r]);
                    if (m == 0) continue;
                    uint32_t at = 0;
                    if (lane_id() == 0) at = atomicAdd(&sh.nout, uint32_t(__popcll(m)));
                    at = __shfl(at, 0, kWave);
                    if (ok[r])
                        sh.out[at + __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u))] =
                            e[r];
                }
                __syncthreads();
                if (sh.nout > sh.kOutCap - kBlock * EU) bin_out_flush(&sh, st, ctl, rp, rec0, lim);
            }
            __syncthreads();  // the batch tables are refilled next round
        }
        bin_out_flush(&sh, st, ctl, rp, rec0, lim);
        uint32_t* bh = bg.bh + uint64_t(i) * bg.cmax;
        for (uint32_t t = threadIdx.x; t < bg.C; t += kBlock)
            if (sh.hist[t]) atomicAdd(&bh[t], sh.hist[t]);
        __syncthreads();
    }
}

template <uint32_t B>
struct BinScatterShared {
    int32_t mode[64];
    unsigned long long rbase[64];
    uint32_t cb[kBinMaxC + 1];
    uint16_t gs[kBinMaxG + 1];
    uint32_t boff[kBinMaxC];  // region offset of each coarse bin within the instance's slice
    uint32_t hist[kBinMaxC];  // the batch's records per bin, then each bin's first slot in `sorted` (zero between batches)
    uint32_t wofs[kBinMaxC];  // slice position of sorted slot j of bin c: wofs[c] + j (mod 2^32)
    uint32_t sorted[B];       // the batch's records grouped by bin
    uint32_t tmp[kWavesPerBlock + 1];
    uint32_t bad;             // a reservation passed its bin's count (cannot happen)
    unsigned long long nbatch;
};

// Per binned instance: the records (expansion order) are streamed in equal
// slices over the grid, B = U * kBlock at a time.  A batch is counting-sorted
// by coarse bin in LDS (the rank of each record within its bin comes from the
// histogram's atomic), each bin holding records reserves its run with one
// atomic on the bin's cursor, and the block writes the sorted batch out:
// consecutive threads write consecutive slots of a run.
template <uint32_t U>
__global__ void __launch_bounds__(kBlock) k_bin_scatter(const Inst* insts, InstState* states, Control* ctl,
                                                        uint32_t I, uint32_t s, uint32_t round, BinGeo bg) {
    constexpr uint32_t B = U * kBlock;
    static_assert(B <= 65536 && kBinMaxC <= 65536, "bin and rank are packed in 16 bits each");
    __shared__ BinScatterShared<B> sh;
    const uint32_t par = s & 1u;
    const uint32_t C = bg.C;
    if (threadIdx.x < I) {
        sh.mode[threadIdx.x] = states[threadIdx.x].bflag[round & 1u];
        sh.rbase[threadIdx.x] = states[threadIdx.x].rbase[par];
    }
    for (uint32_t t = threadIdx.x; t <= C; t += kBlock) sh.cb[t] = bg.cbw[t];
    for (uint32_t t = threadIdx.x; t <= bg.ng; t += kBlock) sh.gs[t] = bg.gsel[t];
    for (uint32_t t = threadIdx.x; t < C; t += kBlock) sh.hist[t] = 0;
    if (threadIdx.x == 0) {
        sh.bad = 0;
        sh.nbatch = 0;
    }
    __syncthreads();
    // bins c0 = 2 t and c0 + 1 belong to thread t in the per-batch scan (C <= 2 kBlock)
    const uint32_t c0 = 2 * threadIdx.x, c1 = c0 + 1;
    for (uint32_t i = 0; i < I; ++i) {
        if (!sh.mode[i]) continue;
        const uint32_t* bh = bg.bh + uint64_t(i) * bg.cmax;
        uint32_t* bcur = bg.bc + uint64_t(i) * bg.cmax;
        const uint64_t rb = sh.rbase[i];
        // region offsets: exclusive prefix of the counts (block 0 publishes them)
        uint32_t carry = 0;
        for (uint32_t q0 = 0; q0 < C; q0 += kBlock) {
            const uint32_t c = q0 + threadIdx.x;
            const uint32_t x = c < C ? bh[c] : 0u;
            uint32_t tot;
            const uint32_t ex = block_excl_scan(x, &tot, sh.tmp);
            if (c < C) {
                sh.boff[c] = carry + ex;
                if (blockIdx.x == 0) bg.bb[uint64_t(i) * bg.cmax + c] = rb + carry + ex;
            }
            carry += tot;
        }
        __syncthreads();
        const uint64_t R = states[i].r0cnt[round & 1u][0];
        const uint64_t lo = R * blockIdx.x / gridDim.x, hi = R * (blockIdx.x + 1) / gridDim.x;
        const uint32_t* rec0 = bg.rec2 + rb;
        uint32_t* rec1 = bg.rec1 + rb;
        for (uint64_t t0 = lo; t0 < hi; t0 += B) {
            const uint32_t nb = hi - t0 < B ? uint32_t(hi - t0) : B;
            uint32_t e[U], cr[U];  // record; its bin (high 16 bits) and rank in the bin (low 16)
#pragma unroll
            for (uint32_t r = 0; r < U; ++r) {
                const uint32_t j = r * kBlock + threadIdx.x;
                e[r] = j < nb ? rec0[t0 + j] : 0u;
            }
#pragma unroll
            for (uint32_t r = 0; r < U; ++r) {
                const uint32_t j = r * kBlock + threadIdx.x;
                cr[r] = 0;
                if (j < nb) {
                    const uint32_t c = coarse_of(sh.cb, sh.gs, bg.gsh, e[r]);
                    cr[r] = (c << 16) | atomicAdd(&sh.hist[c], 1u);
                }
            }
            __syncthreads();
            {
                // the batch's first slot per bin, and one reservation per bin
                // that holds records (both of a thread's go out together)
                const uint32_t x0 = c0 < C ? sh.hist[c0] : 0u, x1 = c1 < C ? sh.hist[c1] : 0u;
                uint32_t tot;
                const uint32_t ex = block_excl_scan(x0 + x1, &tot, sh.tmp);
                uint32_t g0 = 0, g1 = 0;
                if (x0) g0 = atomicAdd(&bcur[c0], x0);
                if (x1) g1 = atomicAdd(&bcur[c1], x1);
                if (x0) {
                    if (uint64_t(g0) + x0 > bh[c0]) sh.bad = 1;
                    sh.hist[c0] = ex;
                    sh.wofs[c0] = sh.boff[c0] + g0 - ex;
                }
                if (x1) {
                    if (uint64_t(g1) + x1 > bh[c1]) sh.bad = 1;
                    sh.hist[c1] = ex + x0;
                    sh.wofs[c1] = sh.boff[c1] + g1 - (ex + x0);
                }
            }
            __syncthreads();
#pragma unroll
            for (uint32_t r = 0; r < U; ++r)
                if (r * kBlock + threadIdx.x < nb) sh.sorted[sh.hist[cr[r] >> 16] + (cr[r] & 0xFFFFu)] = e[r];
            __syncthreads();
            if (c0 < C) sh.hist[c0] = 0;
            if (c1 < C) sh.hist[c1] = 0;
            if (sh.bad) {
                // a run would pass its bin's region (the expand counted these
                // very records): nothing more is written, the error is raised
                if (threadIdx.x == 0) atomicCAS(&ctl->error, 0, int32_t(kErrBin));
                break;
            }
            for (uint32_t j = threadIdx.x; j < nb; j += kBlock) {
                const uint32_t x = sh.sorted[j];
                rec1[sh.wofs[coarse_of(sh.cb, sh.gs, bg.gsh, x)] + j] = x;
            }
            if (threadIdx.x == 0) sh.nbatch++;
            __syncthreads();  // `sorted` and `wofs` are refilled by the next batch
        }
        if (sh.bad) break;
    }
    if (threadIdx.x == 0 && sh.nbatch) atomicAdd(&ctl->dbg[0], sh.nbatch);
}

constexpr uint32_t kBinMaxVb = 4096;   // vertices per fine bin at most
constexpr uint32_t kBinFineBatch = 128;  // fine bins a block lists at a time
// Removals listed in LDS before they are pushed: room for two sweep rounds
// of 512 vertices at 9216 counter words (3 blocks per CU still fit in LDS),
// one round for the other sizes.
constexpr uint32_t kBinRemCap = 1024;
template <uint32_t DW>
constexpr uint32_t bin_rem_cap() { return DW == kBinDeltaWords3 ? 1536u : kBinRemCap; }

template <uint32_t DW>
struct BinApplyShared {
    int32_t mode[64];
    int32_t level[64];
    __attribute__((aligned(16))) uint32_t delta[DW];  // records per degree word of the fine bin (zero between uses)
    static constexpr uint32_t kRem = bin_rem_cap<DW>();
    uint32_t rem[kRem];
    uint32_t nrem;
    uint32_t tmp[kWavesPerBlock + 1];
    uint4 fdesc[kBinFineBatch];       // listed fine bins: records [x, y), first word z, words w
    unsigned long long foff[kBinFineBatch];  // their records' buffer offset from rec1
    uint32_t nlist;
    uint32_t claim[kBinMaxVb / 32];  // vertices of the fine bin already updated (record-driven path; zero between uses)
    unsigned long long base;
    unsigned long long ent;  // entries of the removals this block pushed for the current instance
    uint32_t removed;
};

// Pushes the listed removals as chunks for the cascade.  All threads call.
template <class SH>
__device__ __forceinline__ void bin_push_removed(const Inst& in, InstState* st, Control* ctl, SH* sh,
                                                 uint32_t par, uint32_t rp) {
    const uint32_t nr = sh->nrem;
    for (uint32_t q0 = 0; q0 < nr; q0 += kBlock) {
        const uint32_t q = q0 + threadIdx.x;
        const bool has = q < nr;
        const uint32_t v = has ? sh->rem[q] : 0u;
        if (has) atomicAdd(&sh->ent, (unsigned long long)(in.off[v + 1] - in.off[v]));
        block_defer_push(in, defer_list(in, par), &st->pbr[rp][0], ctl, has, v, sh->tmp, &sh->base);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        sh->removed += nr;
        sh->nrem = 0;
    }
    __syncthreads();
}

// Vertex uu of the fine bin starting at vertex fv0: subtracts its record
// counts from its row (its fine bin owns the row), clears them, stores the new
// count of layers at or above thr and removes the vertex (index k) when
// fewer than lmd are left; returns whether it did, with the vertex in *u.
template <class SH>
__device__ __forceinline__ bool bin_row_update(const Inst& in, SH* sh, uint32_t uu, uint32_t fv0,
                                               uint32_t thr, int32_t k, bool quad, uint32_t* u) {
    const uint32_t L = in.L;
    uint32_t* dl = &sh->delta[uu * L];
    *u = fv0 + uu;
    uint32_t* row = rec_of(in, *u);
    uint32_t c = 0;
    if (quad) {
        for (uint32_t w = 0; w < L; w += 4) {
            uint4 x = *reinterpret_cast<const uint4*>(row + w);
            const uint4 d = *reinterpret_cast<const uint4*>(dl + w);
            // d <= x: each record is a removed neighbour the word still counts
            x.x -= d.x;
            x.y -= d.y;
            x.z -= d.z;
            x.w -= d.w;
            *reinterpret_cast<uint4*>(row + w) = x;
            *reinterpret_cast<uint4*>(dl + w) = make_uint4(0u, 0u, 0u, 0u);
            c += uint32_t(x.x >= thr) + uint32_t(x.y >= thr) + uint32_t(x.z >= thr) + uint32_t(x.w >= thr);
        }
    } else {
        for (uint32_t l = 0; l < L; ++l) {
            const uint32_t x = row[l] - dl[l];
            row[l] = x;
            dl[l] = 0;
            c += x >= thr;
        }
    }
    *cnt_of(in, *u) = int32_t(c);
    if (c >= in.lmd) return false;  // the vertex was alive (at least lmd layers) before these records
    *idx_of(in, *u) = k;
    clear_alive(in, *u);
    return true;
}

// bin_row_update for L = 8 with the row already loaded (x0, x1): the sweep
// issues the loads of two rows before either is used.
template <class SH>
__device__ __forceinline__ bool bin_row_update8(const Inst& in, SH* sh, uint32_t uu, uint32_t fv0, uint32_t thr,
                                                int32_t k, uint4 x0, uint4 x1) {
    uint32_t* dl = &sh->delta[uu * 8];
    const uint32_t u = fv0 + uu;
    const uint4 d0 = *reinterpret_cast<const uint4*>(dl), d1 = *reinterpret_cast<const uint4*>(dl + 4);
    // d <= x: each record is a removed neighbour the word still counts
    x0.x -= d0.x;
    x0.y -= d0.y;
    x0.z -= d0.z;
    x0.w -= d0.w;
    x1.x -= d1.x;
    x1.y -= d1.y;
    x1.z -= d1.z;
    x1.w -= d1.w;
    uint4* row = reinterpret_cast<uint4*>(rec_of(in, u));
    row[0] = x0;
    row[1] = x1;
    *reinterpret_cast<uint4*>(dl) = make_uint4(0u, 0u, 0u, 0u);
    *reinterpret_cast<uint4*>(dl + 4) = make_uint4(0u, 0u, 0u, 0u);
    const uint32_t c = uint32_t(x0.x >= thr) + uint32_t(x0.y >= thr) + uint32_t(x0.z >= thr) + uint32_t(x0.w >= thr) +
                       uint32_t(x1.x >= thr) + uint32_t(x1.y >= thr) + uint32_t(x1.z >= thr) + uint32_t(x1.w >= thr);
    *cnt_of(in, u) = int32_t(c);
    if (c >= in.lmd) return false;  // the vertex was alive (at least lmd layers) before these records
    *idx_of(in, u) = k;
    clear_alive(in, u);
    return true;
}

// Any record counted for vertex uu of the fine bin (L = 8).
template <class SH>
__device__ __forceinline__ bool bin_touched8(const SH* sh, uint32_t uu) {
    const uint4 x = *reinterpret_cast<const uint4*>(&sh->delta[uu * 8]);
    const uint4 y = *reinterpret_cast<const uint4*>(&sh->delta[uu * 8 + 4]);
    return (x.x | x.y | x.z | x.w | y.x | y.y | y.z | y.w) != 0;
}

// The records [a, b) of one fine bin (degree words [fw0, fw0 + nw)): counted
// per word in LDS, then a sweep over the bin's vertices: a vertex with any
// count subtracts them from its row (the fine bin owns these rows), clears them,
// and recounts its layers at or above the threshold; with fewer than lmd it
// is removed.  The counters are zero again afterwards.  All threads call.
template <class SH>
__device__ __forceinline__ void bin_apply_fine(const Inst& in, InstState* st, Control* ctl, SH* sh,
                                               const uint32_t* recf, uint32_t a, uint32_t b,
                                               uint32_t fw0, uint32_t nw, uint32_t par, uint32_t rp, int32_t k,
                                               uint32_t claim_q) {
    const uint32_t L = in.L, thr = uint32_t(k) + 1u;
    const uint32_t fv0 = fw0 / L, nv = nw / L;
    const bool quad = (L & 3u) == 0;
    const bool few = uint64_t(b - a) * 4 < uint64_t(nv) * claim_q;
    // every record lies in [fw0, fw0 + nw) (the passes before placed it
    // there); one that did not would be skipped, never index LDS out of range
    if (few && b - a <= kBlock * kBinUnroll) {
        // few records, all held in registers from the count to the claims
        // (no second load): each vertex is claimed by its first record (one
        // LDS bit) and updated once
        uint32_t e[kBinUnroll];
#pragma unroll
        for (uint32_t r = 0; r < kBinUnroll; ++r) {
            const uint32_t t = a + r * kBlock + threadIdx.x;
            e[r] = t < b ? recf[t] : 0xFFFFFFFFu;
        }
#pragma unroll
        for (uint32_t r = 0; r < kBinUnroll; ++r)
            if (e[r] - fw0 < nw) atomicAdd(&sh->delta[e[r] - fw0], 1u);
        __syncthreads();
        for (uint32_t r = 0; r < kBinUnroll && a + r * kBlock < b; ++r) {
            bool rem = false;
            uint32_t u = 0;
            if (e[r] - fw0 < nw) {
                const uint32_t uu = div_L(e[r] - fw0, L, in.lmagic);
                const uint32_t bit = 1u << (uu & 31u);
                if ((atomicOr(&sh->claim[uu >> 5], bit) & bit) == 0) rem = bin_row_update(in, sh, uu, fv0, thr, k, quad, &u);
            }
            if (rem) sh->rem[atomicAdd(&sh->nrem, 1u)] = u;
            __syncthreads();
            if (sh->nrem > SH::kRem - kBlock) bin_push_removed(in, st, ctl, sh, par, rp);
        }
        for (uint32_t w = threadIdx.x; w < (nv + 31) / 32; w += kBlock) sh->claim[w] = 0;
        __syncthreads();
        return;
    }
    for (uint32_t t0 = a + threadIdx.x; t0 < b; t0 += kBlock * kBinUnroll) {
        uint32_t e[kBinUnroll];
#pragma unroll
        for (uint32_t r = 0; r < kBinUnroll; ++r) {
            const uint32_t t = t0 + r * kBlock;
            e[r] = t < b ? recf[t] : 0xFFFFFFFFu;
        }
#pragma unroll
        for (uint32_t r = 0; r < kBinUnroll; ++r)
            if (e[r] - fw0 < nw) atomicAdd(&sh->delta[e[r] - fw0], 1u);
    }
    __syncthreads();
    if (few) {
        // few records: each vertex is claimed by its first record (one LDS
        // bit) and updated once; work follows the records, not the vertices
        for (uint32_t t0 = a; t0 < b; t0 += kBlock) {
            const uint32_t t = t0 + threadIdx.x;
            bool rem = false;
            uint32_t u = 0;
            const uint32_t e = t < b ? recf[t] : 0xFFFFFFFFu;
            if (e - fw0 < nw) {
                const uint32_t uu = div_L(e - fw0, L, in.lmagic);
                const uint32_t bit = 1u << (uu & 31u);
                if ((atomicOr(&sh->claim[uu >> 5], bit) & bit) == 0) rem = bin_row_update(in, sh, uu, fv0, thr, k, quad, &u);
            }
            if (rem) sh->rem[atomicAdd(&sh->nrem, 1u)] = u;
            __syncthreads();
            if (sh->nrem > SH::kRem - kBlock) bin_push_removed(in, st, ctl, sh, par, rp);
        }
        for (uint32_t w = threadIdx.x; w < (nv + 31) / 32; w += kBlock) sh->claim[w] = 0;
        __syncthreads();
        return;
    }
    if (L == 8 && SH::kRem >= 3 * kBlock) {
        // two vertices per thread and round, their rows' loads in flight together
        for (uint32_t ub = 0; ub < nv; ub += 2 * kBlock) {
            const uint32_t u0 = ub + threadIdx.x, u1 = u0 + kBlock;
            const bool t0 = u0 < nv && bin_touched8(sh, u0);
            const bool t1 = u1 < nv && bin_touched8(sh, u1);
            uint4 a0 = make_uint4(0u, 0u, 0u, 0u), a1 = a0, b0 = a0, b1 = a0;
            if (t0) {
                const uint4* r = reinterpret_cast<const uint4*>(rec_of(in, fv0 + u0));
                a0 = r[0];
                a1 = r[1];
            }
            if (t1) {
                const uint4* r = reinterpret_cast<const uint4*>(rec_of(in, fv0 + u1));
                b0 = r[0];
                b1 = r[1];
            }
            const bool r0 = t0 && bin_row_update8(in, sh, u0, fv0, thr, k, a0, a1);
            const bool r1 = t1 && bin_row_update8(in, sh, u1, fv0, thr, k, b0, b1);
            if (r0) sh->rem[atomicAdd(&sh->nrem, 1u)] = fv0 + u0;
            if (r1) sh->rem[atomicAdd(&sh->nrem, 1u)] = fv0 + u1;
            __syncthreads();
            if (sh->nrem > SH::kRem - 2 * kBlock) bin_push_removed(in, st, ctl, sh, par, rp);
        }
        return;
    }
    for (uint32_t ub = 0; ub < nv; ub += kBlock) {
        const uint32_t uu = ub + threadIdx.x;
        bool rem = false;
        uint32_t u = 0;
        if (uu < nv) {
            const uint32_t* dl = &sh->delta[uu * L];
            uint32_t any = 0;
            if (quad) {
                for (uint32_t w = 0; w < L; w += 4) {
                    const uint4 x = *reinterpret_cast<const uint4*>(dl + w);
                    any |= x.x | x.y | x.z | x.w;
                }
            } else {
                for (uint32_t l = 0; l < L; ++l) any |= dl[l];
            }
            if (any) rem = bin_row_update(in, sh, uu, fv0, thr, k, quad, &u);
        }
        if (rem) sh->rem[atomicAdd(&sh->nrem, 1u)] = u;
        __syncthreads();
        if (sh->nrem > SH::kRem - kBlock) bin_push_removed(in, st, ctl, sh, par, rp);
    }
}

// Grouping of the coarse bins' records by fine bin: a batch of the region's
// records counting-sorted in LDS, written out in runs by consecutive threads.
template <uint32_t FU>
struct BinGroupShared {
    int32_t mode[64];
    uint32_t first[65];                 // tasks of the instances before i
    uint32_t fstart[kBinMaxF + 1];
    uint32_t fcur[kBinMaxF];
    uint32_t fhist[kBinMaxF];  // a batch's records per fine bin, then its first slot (zero between batches)
    uint32_t fwo[kBinMaxF];    // region position of a batch's sorted slot j of fine bin f: fwo[f] + j (mod 2^32)
    uint32_t buf[FU * kBlock];
    uint32_t tmp[kWavesPerBlock + 1];
};

// Tasks (instance, coarse bin) of the binned instances, dealt round robin
// over the grid: a coarse bin of several fine bins has its region's records
// grouped by fine bin into rec2 (same region); every fine bin's record range
// is published in fse (a coarse bin of one fine bin keeps its records in
// rec1).  The bin's counts are cleared for the next round.
template <uint32_t FU>
__global__ void __launch_bounds__(kBlock) k_bin_group(InstState* states, Control* ctl, uint32_t I, uint32_t round,
                                                      BinGeo bg) {
    constexpr uint32_t FB = FU * kBlock;
    static_assert(FB <= 65536 && kBinMaxF <= 65536, "fine bin and rank are packed in 16 bits each");
    __shared__ BinGroupShared<FU> sh;
    const uint32_t rp = round & 1u;
    if (threadIdx.x < I) sh.mode[threadIdx.x] = states[threadIdx.x].bflag[rp];
    // the round's record count was read by the scatter; the next use of this
    // slot (two rounds on) counts from zero
    if (blockIdx.x == 0 && threadIdx.x < I && states[threadIdx.x].bflag[rp]) states[threadIdx.x].r0cnt[rp][0] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t i = 0; i < I; ++i) {
            sh.first[i] = t;
            if (sh.mode[i]) t += bg.C;
        }
        sh.first[I] = t;
    }
    __syncthreads();
    const uint32_t ntasks = sh.first[I];
    const uint32_t vbw = bg.vbw;
    for (uint32_t t = blockIdx.x; t < ntasks; t += gridDim.x) {
        uint32_t i = 0;
        while (!(sh.mode[i] && t < sh.first[i] + bg.C)) ++i;  // t < ntasks: some binned instance holds it
        const uint32_t c = t - sh.first[i];
        const uint64_t q = uint64_t(i) * bg.cmax + c;
        const uint32_t nrec = bg.bc[q];
        const uint64_t base = bg.bb[q];
        const uint32_t w0 = bg.cbw[c], w1 = bg.cbw[c + 1];
        const uint32_t nf = (w1 - w0 + vbw - 1) / vbw;  // <= kBinMaxF (the host sizes coarse bins so)
        uint2* fse = bg.fse + uint64_t(i) * bg.nfine + bg.ffirst[c];
        if (nf == 1 || nrec == 0) {
            for (uint32_t f = threadIdx.x; f < nf; f += kBlock) fse[f] = make_uint2(0u, nf == 1 ? nrec : 0u);
        } else {
            for (uint32_t f = threadIdx.x; f <= nf; f += kBlock) sh.fstart[f] = 0;
            for (uint32_t f = threadIdx.x; f < nf; f += kBlock) sh.fhist[f] = 0;
            __syncthreads();
            const uint32_t* r1 = bg.rec1 + base;
            for (uint32_t t0 = threadIdx.x; t0 < nrec; t0 += kBlock * kBinUnroll) {
                uint32_t e[kBinUnroll];
#pragma unroll
                for (uint32_t r = 0; r < kBinUnroll; ++r) {
                    const uint32_t j = t0 + r * kBlock;
                    e[r] = j < nrec ? r1[j] : 0xFFFFFFFFu;
                }
#pragma unroll
                for (uint32_t r = 0; r < kBinUnroll; ++r) {
                    if (e[r] == 0xFFFFFFFFu) continue;
                    if (e[r] >= w0 && e[r] < w1) atomicAdd(&sh.fstart[(e[r] - w0) / vbw], 1u);
                    else atomicCAS(&ctl->error, 0, int32_t(kErrBin));  // cannot happen: the scatter binned it here
                }
            }
            __syncthreads();
            {
                const uint32_t x = threadIdx.x < nf ? sh.fstart[threadIdx.x] : 0u;
                uint32_t tot;
                const uint32_t ex = block_excl_scan(x, &tot, sh.tmp);
                if (threadIdx.x < nf) {
                    sh.fstart[threadIdx.x] = ex;
                    sh.fcur[threadIdx.x] = ex;
                }
                if (threadIdx.x == 0) sh.fstart[nf] = tot;
            }
            __syncthreads();
            uint32_t* r2 = bg.rec2 + base;
            for (uint32_t t0 = 0; t0 < nrec; t0 += FB) {
                const uint32_t nb = nrec - t0 < FB ? nrec - t0 : FB;
                uint32_t e[FU], fr[FU];  // record; its fine bin (high 16 bits) and rank in the bin (low 16)
#pragma unroll
                for (uint32_t r = 0; r < FU; ++r) {
                    const uint32_t j = r * kBlock + threadIdx.x;
                    e[r] = j < nb ? r1[t0 + j] : 0xFFFFFFFFu;
                }
#pragma unroll
                for (uint32_t r = 0; r < FU; ++r) {
                    fr[r] = 0xFFFFFFFFu;
                    if (e[r] >= w0 && e[r] < w1) {
                        const uint32_t f = (e[r] - w0) / vbw;
                        fr[r] = (f << 16) | atomicAdd(&sh.fhist[f], 1u);
                    }
                }
                __syncthreads();
                uint32_t tot;
                {
                    const uint32_t x = threadIdx.x < nf ? sh.fhist[threadIdx.x] : 0u;
                    const uint32_t ex = block_excl_scan(x, &tot, sh.tmp);
                    if (threadIdx.x < nf) {
                        sh.fhist[threadIdx.x] = ex;
                        sh.fwo[threadIdx.x] = sh.fcur[threadIdx.x] - ex;
                        sh.fcur[threadIdx.x] += x;
                    }
                }
                __syncthreads();
#pragma unroll
                for (uint32_t r = 0; r < FU; ++r)
                    if (fr[r] != 0xFFFFFFFFu) sh.buf[sh.fhist[fr[r] >> 16] + (fr[r] & 0xFFFFu)] = e[r];
                __syncthreads();
                if (threadIdx.x < nf) sh.fhist[threadIdx.x] = 0;
                for (uint32_t j = threadIdx.x; j < tot; j += kBlock) {
                    const uint32_t x = sh.buf[j];
                    r2[sh.fwo[(x - w0) / vbw] + j] = x;
                }
                __syncthreads();  // `buf` and `fwo` are refilled by the next batch
            }
            for (uint32_t f = threadIdx.x; f < nf; f += kBlock) fse[f] = make_uint2(sh.fstart[f], sh.fstart[f + 1]);
        }
        if (threadIdx.x == 0) {
            bg.bh[q] = 0;
            bg.bc[q] = 0;
            if (nrec) atomicAdd(&states[i].brec[0], (unsigned long long)nrec);
        }
        __syncthreads();
    }
}

// The fine bins of the binned instances, dealt interleaved over the grid
// (block b takes fine bins b, b + G, ... of each instance, so every block
// gets a share of the dense and the sparse ranges): each fine bin with
// records is applied in LDS (bin_apply_fine), its removals pushed as chunks
// of the round's list.
template <uint32_t DW>
__global__ void __launch_bounds__(kBlock) k_bin_fine(const Inst* insts, InstState* states, Control* ctl, uint32_t I,
                                                     uint32_t s, uint32_t round, BinGeo bg) {
    __shared__ BinApplyShared<DW> sh;
    const uint32_t par = s & 1u, rp = round & 1u;
    if (threadIdx.x < I) {
        sh.mode[threadIdx.x] = states[threadIdx.x].bflag[rp];
        sh.level[threadIdx.x] = states[threadIdx.x].dk[par];
    }
    for (uint32_t t = threadIdx.x; t < DW; t += kBlock) sh.delta[t] = 0;
    for (uint32_t t = threadIdx.x; t < kBinMaxVb / 32; t += kBlock) sh.claim[t] = 0;
    if (threadIdx.x == 0) {
        sh.removed = 0;
        sh.ent = 0;
        sh.nrem = 0;
    }
    __syncthreads();
    const uint32_t G = gridDim.x, vbw = bg.vbw;
    const uint64_t rcap = bg.rec2 - bg.rec1;
    for (uint32_t i = 0; i < I; ++i) {
        if (!sh.mode[i]) continue;
        const Inst& in = insts[i];
        InstState* st = &states[i];
        const int32_t k = sh.level[i];
        const uint2* fse = bg.fse + uint64_t(i) * bg.nfine;
        for (uint64_t j0 = 0; j0 * G + blockIdx.x < bg.nfine; j0 += kBinFineBatch) {
            if (threadIdx.x == 0) sh.nlist = 0;
            __syncthreads();
            const uint64_t F = (j0 + threadIdx.x) * G + blockIdx.x;
            if (threadIdx.x < kBinFineBatch && F < bg.nfine) {
                const uint2 ab = fse[F];
                if (ab.y > ab.x) {
                    const uint32_t c = bg.fcoarse[F];
                    const uint32_t w0 = bg.cbw[c], w1 = bg.cbw[c + 1];
                    const uint32_t fw0 = w0 + (uint32_t(F) - bg.ffirst[c]) * vbw;
                    const uint32_t at = atomicAdd(&sh.nlist, 1u);
                    sh.fdesc[at] = make_uint4(ab.x, ab.y, fw0, w1 - fw0 < vbw ? w1 - fw0 : vbw);
                    sh.foff[at] = (w1 - w0 <= vbw ? 0ull : rcap) + bg.bb[uint64_t(i) * bg.cmax + c];
                }
            }
            __syncthreads();
            const uint32_t nl = sh.nlist;
            for (uint32_t m = 0; m < nl; ++m) {
                const uint4 d = sh.fdesc[m];
                bin_apply_fine(in, st, ctl, &sh, bg.rec1 + sh.foff[m], d.x, d.y, d.z, d.w, par, rp, k,
                               bg.claim_q);
            }
            __syncthreads();  // the list is refilled next
        }
        bin_push_removed(in, st, ctl, &sh, par, rp);
        if (threadIdx.x == 0) {
            if (sh.removed) atomicAdd(&st->alive[0], 0ull - (unsigned long long)sh.removed);
            if (sh.ent) atomicAdd(&st->pber[rp][0], sh.ent);
            sh.removed = 0;
            sh.ent = 0;
        }
        __syncthreads();
    }
}

// Entry boundaries of the coarse bins: bnd[j] = the first vertex v with
// off[v] >= j * E, j = 0..J (a binary search per j).
__global__ void k_bin_bounds(const int64_t* off, uint32_t n, uint64_t E, uint32_t J, uint32_t* bnd) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > J) return;
    const int64_t want = int64_t(uint64_t(j) * E);
    uint32_t lo = 0, hi = n;  // off[n] = D >= every want < D; want >= D gives n
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (off[mid] >= want) hi = mid;
        else lo = mid + 1;
    }
    bnd[j] = lo;
}

}  // namespace
}  // namespace hg
