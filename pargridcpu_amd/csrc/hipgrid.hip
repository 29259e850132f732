// hipgrid.hip -- C ABI of the HipGrid FirmCore engine (include/hipgrid.h).
//
// Device memory layout of one graph (n vertices, L layers, D directed
// entries over all layers):
//   (internal vertex ids: caller's ids sorted by total degree, descending)
//   off     int64 [n+1]       vertex-major adjacency offsets
//   adj     uint32 [D]        entry u*L + l (neighbour u in layer l)
//   adjz    uint8 [D]         z(u) of each entry's target
//   deg0    uint32 [n][L]     initial layer degrees
//   z       uint8 [n]         number of layers in which the vertex has an edge
//   inv     uint32 [n]        caller's id -> internal id
//   chunks  uint64 [nchunks]  (v, chunk#) list covering adj in kChunk pieces
//   out     int32 [L][n]      core index in the caller's ids (the result)
// and per instance lmd = 1..L (all allocated once, reused by every call):
//   deg uint32 [n][L], cnt int32 [n], idx int32 [n] (internal ids),
//   per-block alive list and removal ring uint32 [G*cap] each (block b of
//   the fixed grid owns vertices [b*cap, (b+1)*cap)), two deferred chunk
//   lists uint64 [pcap] (pcap = sum over v of ceil(deg(v) / kMinDefer)).
// The graph stays resident between calls; hg_decompose touches no host
// memory except a small control word per batch of levels.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "hipgrid.h"
#include "hipgrid_common.h"
#include "hipgrid_kernels.h"
#include "hipgrid_bin_kernels.h"

using namespace hg;
using hgc::dalloc;
using hgc::env_int;
using hgc::fail;

struct hg_graph {
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 0;
    uint64_t wall_khz = 100000;
    uint32_t n = 0, L = 0, kcap = 0;
    uint32_t G = 0, cap = 0;  // level-kernel grid and vertices per block
    uint32_t max_nz = 0;      // most layers in which one vertex has an edge
    uint64_t D = 0, nchunks = 0, pcap = 0;
    int64_t device_bytes = 0;
    // graph
    int64_t* off = nullptr;
    uint32_t* adj = nullptr;
    uint8_t* adjz = nullptr;
    uint8_t* adjh = nullptr;  // [L][D] core-index bounds per entry (or null)
    uint32_t* deg0 = nullptr;
    uint32_t* inv = nullptr;
    uint8_t* z = nullptr;
    uint32_t* alive = nullptr;  // per instance ceil(n/32) words
    uint32_t H = 0;             // vertices (lowest internal ids) pulled by a wavefront each
    uint32_t pull = 0;          // level 0 by pull (L <= kPullMaxL)
    int32_t* out = nullptr;
    uint64_t* chunks = nullptr;
    // instances
    uint32_t* deg = nullptr;
    int32_t* cnt = nullptr;
    int32_t* idx = nullptr;
    uint32_t* lists = nullptr;
    uint32_t* rings = nullptr;
    uint32_t* blk = nullptr;
    uint64_t* pl = nullptr;
    uint64_t* d_ihist = nullptr;  // [L][kcap] core index histogram
    Inst* d_insts = nullptr;
    InstState* d_states = nullptr;
    Control* d_ctl = nullptr;
    Control* h_ctl = nullptr;  // pinned
    std::vector<Inst> insts;
    // densest
    uint64_t* hist = nullptr;
    double* d_density = nullptr;
    int32_t* d_sel = nullptr;
    int64_t* d_cell_len = nullptr;
    double* d_pow = nullptr;
    // results
    bool decomposed = false;
    std::vector<int64_t> h_count;    // [L][kcap]: vertices with core index k
    std::vector<int32_t> h_kmax;     // [L]
    bool densest_done = false;
    hg_wds wds{};
    std::vector<double> h_density;   // [L][kcap]
    std::vector<int32_t> h_sel;
    hg_timing timing{};
    // host-side scratch of the per-call uploads
    std::vector<InstState> h_states;
    Control h_ctl_init{};
    std::vector<double> h_pow;
    std::vector<int64_t> h_len;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // phase timing (hg_set_phase_timing): events before each step's scan,
    // between its scan and cascade, and after the last cascade
    bool phase_timing = false;
    std::vector<hipEvent_t> pev;
    // row compaction (k_cmp_*): per instance two row buffers used in turn,
    // plus shared scratch for the counts, prefix sums and chunk list
    struct CmpBuf {
        int64_t* off[2] = {nullptr, nullptr};
        uint32_t* adj[2] = {nullptr, nullptr};
        uint64_t adj_cap[2] = {0, 0};
    };
    std::vector<CmpBuf> cmp;
    uint64_t* c_len = nullptr;       // [n+1] new row lengths, then cursors
    uint64_t* c_nch = nullptr;       // [n+1] chunks per current row
    uint64_t* c_item_off = nullptr;  // [n+1]
    uint64_t* c_items = nullptr;
    uint64_t c_items_cap = 0;
    uint8_t* c_temp = nullptr;
    size_t c_temp_bytes = 0;
    unsigned long long* c_words = nullptr;
    uint64_t compact_min = 0;  // compaction for graphs of at least this many entries (0: never)
    uint64_t compact_ratio = 2;  // compact when alive entries <= rows' entries / this
    int64_t cmp_bytes = 0;
    std::vector<Inst> cur;     // the instances' current rows during hg_decompose
    int64_t compactions = 0;   // in the last hg_decompose
    // binned apply of dense removal rounds (k_bin_*): coarse bin table, per-bin
    // counts / cursors / region starts per instance, two record buffers
    bool bin_on = false;
    uint32_t bin_steps = 0;  // steps 1..bin_steps may bin (and step 0 with bin_level0)
    uint32_t bin_rounds = 0; // removal rounds 0..bin_rounds-1 of such a step may bin
    uint32_t bin_eu = 4;     // entries per thread in flight in k_bin_expand (4 or 8)
    uint32_t bin_xg = 4;     // k_bin_expand blocks per CU
    bool bin_level0 = true;  // step 0 (level 0's pushed frontier) may bin too
    uint32_t bin_sg = 2, bin_ag = 3;  // k_bin_scatter / k_bin_group and k_bin_fine blocks per CU
    BinGeo bgeo{};
    uint32_t* bin_cbw = nullptr;
    uint32_t* bin_bh = nullptr;
    uint32_t* bin_bc = nullptr;
    uint64_t* bin_bb = nullptr;
    uint32_t* bin_rec = nullptr;  // rec1 then rec2
    uint16_t* bin_gsel = nullptr;
    uint32_t* bin_ffirst = nullptr;  // [C+1] first fine bin of each coarse bin
    uint16_t* bin_fcoarse = nullptr; // [nfine] coarse bin of each fine bin
    uint2* bin_fse = nullptr;        // [L][nfine] fine bins' record ranges
};

namespace {

void release(hg_graph* g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    (void)hipStreamSynchronize(g->stream);  // nothing queued may still use the memory
    void* ptrs[] = {g->off, g->adj, g->adjz, g->adjh, g->deg0, g->alive, g->inv, g->z, g->out, g->chunks, g->deg, g->cnt, g->idx, g->lists, g->rings, g->blk, g->pl,
                    g->d_ihist, g->d_insts, g->d_states, g->d_ctl, g->hist, g->d_density, g->d_sel, g->d_cell_len,
                    g->d_pow};
    for (void* p : ptrs)
        if (p) hgc::dfree(p);
    if (g->h_ctl) (void)hipHostFree(g->h_ctl);
    for (auto& e : g->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : g->pev)
        if (e) (void)hipEventDestroy(e);
    for (auto& c : g->cmp)
        for (int b = 0; b < 2; ++b) {
            if (c.off[b]) hgc::dfree(c.off[b]);
            if (c.adj[b]) hgc::dfree(c.adj[b]);
        }
    void* bptrs[] = {g->bin_cbw, g->bin_bh, g->bin_bc, g->bin_bb, g->bin_rec, g->bin_gsel, g->bin_ffirst,
                     g->bin_fcoarse, g->bin_fse};
    for (void* p : bptrs)
        if (p) hgc::dfree(p);
    void* cptrs[] = {g->c_len, g->c_nch, g->c_item_off, g->c_items, g->c_temp, g->c_words};
    for (void* p : cptrs)
        if (p) hgc::dfree(p);
    delete g;
}

// Device memory that lives only through hg_graph_create; freed once the
// stream has drained (also on an early error return, when copies and
// kernels that use it may still be queued).
struct Scratch {
    hipStream_t stream = nullptr;
    std::vector<void*> ptrs;
    ~Scratch() {
        if (!ptrs.empty()) (void)hipStreamSynchronize(stream);
        for (void* p : ptrs)
            if (p) hgc::dfree(p);
    }
    template <typename T>
    hipError_t alloc(T** p, size_t count) {
        const hipError_t e = dalloc(p, count);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

// Coarse bins of the binned apply: vertex ranges of about D / C_target
// entries each (from the merged offsets: hub ranges are short), none longer
// than kBinMaxF fine bins of vb = (the apply's LDS counter words) / L vertices.  A graph whose
// table would exceed kBinMaxC bins keeps the atomic path everywhere.
int bin_setup(hg_graph* g, Scratch& tmp, hgc::Staging& up) {
    const uint32_t L = g->L, n = g->n;
    // blocks per CU of the scatter (2 or 3) and the apply (2 to 4): their LDS sizes
    g->bin_sg = env_int("HG_BIN_SG", 2) == 3 ? 3u : 2u;
    g->bin_ag = uint32_t(std::max(2, std::min(4, env_int("HG_BIN_AG", 3))));
    const uint32_t dw = g->bin_ag == 4 ? kBinDeltaWords4 : g->bin_ag == 3 ? kBinDeltaWords3 : kBinDeltaWords2;
    const uint32_t vb = std::min<uint32_t>(dw / L, kBinMaxVb);
    const uint64_t vlimit = uint64_t(vb) * kBinMaxF;
    const uint64_t ct = uint64_t(std::max(1, env_int("HG_BIN_C", 512)));
    const uint64_t E = std::max<uint64_t>(1, (g->D + ct - 1) / ct);
    const uint32_t J = uint32_t((g->D + E - 1) / E);
    uint32_t* bnd = nullptr;
    HIP_TRY(tmp.alloc(&bnd, size_t(J) + 1));
    hipLaunchKernelGGL(k_bin_bounds, dim3((J + 1 + 255) / 256), dim3(256), 0, g->stream, g->off, n, E, J, bnd);
    HIP_TRY(hipGetLastError());
    std::vector<uint32_t> hb(size_t(J) + 1);
    HIP_TRY(up.d2h(hb.data(), bnd, sizeof(uint32_t) * hb.size()));
    std::vector<uint32_t> cb{0};
    for (uint32_t j = 1; j <= J; ++j) {
        const uint32_t b = j == J ? n : std::min(hb[j], n);
        while (uint64_t(b) - cb.back() > vlimit) cb.push_back(uint32_t(cb.back() + vlimit));
        if (b > cb.back()) cb.push_back(b);
    }
    const uint32_t C = uint32_t(cb.size() - 1);
    if (C < 1 || C > kBinMaxC) return HG_OK;  // binning stays off for this graph
    std::vector<uint32_t> cbw(cb.size());
    for (size_t c = 0; c < cb.size(); ++c) cbw[c] = cb[c] * L;
    // granule table: word granules of 2^gsh words, at most kBinMaxG of them
    const uint64_t nL = uint64_t(n) * L;
    uint32_t gsh = 0;
    while (((nL + (1ull << gsh) - 1) >> gsh) > kBinMaxG) ++gsh;
    const uint32_t ng = uint32_t((nL + (1ull << gsh) - 1) >> gsh);
    std::vector<uint16_t> gsel(size_t(ng) + 1);
    {
        uint32_t c = 0;
        for (uint32_t q = 0; q < ng; ++q) {
            const uint64_t w = uint64_t(q) << gsh;  // first word of granule q
            while (c + 1 < C && cbw[c + 1] <= w) ++c;
            gsel[q] = uint16_t(c);
        }
        gsel[ng] = uint16_t(C - 1);
    }
    // fine bins: vb vertices each, numbered over all coarse bins
    std::vector<uint32_t> ffirst(size_t(C) + 1, 0);
    for (uint32_t c = 0; c < C; ++c) ffirst[c + 1] = ffirst[c] + (cbw[c + 1] - cbw[c] + vb * L - 1) / (vb * L);
    const uint32_t nfine = ffirst[C];
    std::vector<uint16_t> fcoarse(nfine);
    for (uint32_t c = 0; c < C; ++c)
        for (uint32_t f = ffirst[c]; f < ffirst[c + 1]; ++f) fcoarse[f] = uint16_t(c);
    const uint64_t rcap = hgc::bin_rcap(g->D);
    const size_t tab = size_t(L) * C;
    HIP_TRY(dalloc(&g->bin_cbw, cbw.size()));
    HIP_TRY(dalloc(&g->bin_bh, tab));
    HIP_TRY(dalloc(&g->bin_bc, tab));
    HIP_TRY(dalloc(&g->bin_bb, tab));
    HIP_TRY(dalloc(&g->bin_rec, size_t(2 * rcap)));
    HIP_TRY(dalloc(&g->bin_gsel, gsel.size()));
    HIP_TRY(dalloc(&g->bin_ffirst, ffirst.size()));
    HIP_TRY(dalloc(&g->bin_fcoarse, fcoarse.size()));
    HIP_TRY(dalloc(&g->bin_fse, size_t(L) * nfine));
    HIP_TRY(up.h2d(g->bin_ffirst, ffirst.data(), sizeof(uint32_t) * ffirst.size()));
    HIP_TRY(up.h2d(g->bin_fcoarse, fcoarse.data(), sizeof(uint16_t) * fcoarse.size()));
    HIP_TRY(up.h2d(g->bin_cbw, cbw.data(), sizeof(uint32_t) * cbw.size()));
    HIP_TRY(up.h2d(g->bin_gsel, gsel.data(), sizeof(uint16_t) * gsel.size()));
    HIP_TRY(hipMemsetAsync(g->bin_bh, 0, sizeof(uint32_t) * tab, g->stream));
    HIP_TRY(hipMemsetAsync(g->bin_bc, 0, sizeof(uint32_t) * tab, g->stream));
    BinGeo& bg = g->bgeo;
    bg.cbw = g->bin_cbw;
    bg.bh = g->bin_bh;
    bg.bc = g->bin_bc;
    bg.bb = g->bin_bb;
    bg.rec1 = g->bin_rec;
    bg.rec2 = g->bin_rec + rcap;
    bg.rcap = rcap;
    bg.gsel = g->bin_gsel;
    bg.ng = ng;
    bg.gsh = gsh;
    bg.min_ent = uint64_t(std::max(1, env_int("HG_BIN_MINENT", 1 << 20)));
    bg.C = C;
    bg.cmax = C;
    bg.vbw = vb * L;
    bg.ffirst = g->bin_ffirst;
    bg.fcoarse = g->bin_fcoarse;
    bg.fse = g->bin_fse;
    bg.nfine = nfine;
    // binned when the frontier's entries reach HG_BIN_DENS / 1000 of the
    // alive vertices' degree words, in up to HG_BIN_ROUNDS removal rounds
    // per level (later rounds are sparser).  On the 100M-vertex workload
    // 5 % and three rounds: 380-381 ms per decomposition; 10 % and two
    // rounds 386 ms; 2.5 % and three rounds 395 ms; four rounds 380 ms.
    const char* dv = std::getenv("HG_BIN_DENS");
    bg.dens_num = dv && *dv ? uint32_t(std::max(0, std::atoi(dv))) : 50u;
    bg.dens_den = 1000;
    g->bin_steps = uint32_t(env_int("HG_BIN_STEPS", 64));
    g->bin_rounds = uint32_t(env_int("HG_BIN_ROUNDS", 3));
    g->bin_eu = env_int("HG_BIN_EU", 4) == 8 ? 8u : 4u;
    g->bin_level0 = env_int("HG_BIN_L0", 1) == 1;
    bg.cont = env_int("HG_BIN_CONT", 1) == 1 ? 1u : 0u;
    bg.claim_q = uint32_t(std::max(0, std::min(64, env_int("HG_BIN_CLAIM", 4))));
    g->bin_xg = uint32_t(std::min(8, env_int("HG_BIN_XG", 4)));
    g->bin_on = true;
    return HG_OK;
}

hipError_t sort_by_total(void* temp, size_t& temp_bytes, const uint64_t* tot, uint64_t* tot_s, const uint32_t* ids,
                         uint32_t* perm, uint32_t n, hipStream_t stream) {
    // stable: equal totals keep the caller's order
    return hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, tot, tot_s, ids, perm, int(n), 0, 40, stream);
}

// The caller's per-layer CSR goes up as is (through pinned staging); the
// set-up runs on the device: per-vertex totals, the degree order (a stable
// radix sort, largest total first), inverse permutation, merged offsets and
// chunk offsets (prefix sums), then the relabelled vertex-major adjacency.
// The host reads back a handful of words to size the remaining arrays.
int create_impl(const hg_layer_csr* layers, int32_t L, int64_t n, int32_t offset_bits, void* stream, hg_graph* g) {
    if (!layers) return fail(HG_ERR_ARG, "layers is NULL");
    if (L < 1 || L > 64) return fail(HG_ERR_RANGE, "L must be in [1, 64]");
    if (n < 1) return fail(HG_ERR_ARG, "n must be >= 1");
    if (uint64_t(n) * uint64_t(L) >= (1ull << 32)) return fail(HG_ERR_RANGE, "n * L must be < 2^32");
    if (n >= (1ll << 31) - 1) return fail(HG_ERR_RANGE, "n must be < 2^31 - 1");
    if (offset_bits != 32 && offset_bits != 64) return fail(HG_ERR_ARG, "offset_bits must be 32 or 64");
    for (int l = 0; l < L; ++l)
        if (!layers[l].row_ptr || (!layers[l].col && n > 0)) return fail(HG_ERR_ARG, "layer pointer is NULL");

    HIP_TRY(hipGetDevice(&g->device));
    g->stream = static_cast<hipStream_t>(stream);
    HIP_TRY(hipDeviceGetAttribute(&g->cus, hipDeviceAttributeMultiprocessorCount, g->device));
    int wall = 0;
    if (hipDeviceGetAttribute(&wall, hipDeviceAttributeWallClockRate, g->device) == hipSuccess && wall > 0)
        g->wall_khz = uint64_t(wall);
    g->n = uint32_t(n);
    g->L = uint32_t(L);

    // Entries per layer from the end offsets (the rest of the offsets are
    // checked on the device).
    std::vector<uint64_t> m(static_cast<size_t>(L));
    uint64_t D = 0;
    for (int l = 0; l < L; ++l) {
        int64_t first, last;
        if (offset_bits == 64) {
            first = static_cast<const int64_t*>(layers[l].row_ptr)[0];
            last = static_cast<const int64_t*>(layers[l].row_ptr)[n];
        } else {
            first = static_cast<const int32_t*>(layers[l].row_ptr)[0];
            last = static_cast<const int32_t*>(layers[l].row_ptr)[n];
        }
        if (first != 0) return fail(HG_ERR_ARG, "row_ptr[0] must be 0");
        if (last < 0) return fail(HG_ERR_ARG, "row_ptr must be non-decreasing");
        m[size_t(l)] = uint64_t(last);
        D += uint64_t(last);
        if (D >= (1ull << 40)) return fail(HG_ERR_RANGE, "total directed entries must be < 2^40");
    }
    g->D = D;
    const size_t n1 = size_t(n) + 1;
    const uint32_t grid_n = uint32_t(std::min<uint64_t>((uint64_t(n) + kBlock - 1) / kBlock, uint64_t(g->cus) * 16));

    Scratch tmp;
    tmp.stream = g->stream;
    std::vector<int64_t*> d_rp(static_cast<size_t>(L), nullptr);
    std::vector<uint32_t*> d_col(static_cast<size_t>(L), nullptr);
    int64_t** d_rp_arr = nullptr;
    uint32_t** d_col_arr = nullptr;
    int32_t* d_rp32 = nullptr;
    uint64_t *tot = nullptr, *tot_s = nullptr, *chunk_cnt = nullptr, *coff = nullptr;
    uint32_t *ids = nullptr, *perm = nullptr;
    uint8_t* nzc = nullptr;
    unsigned long long* prep = nullptr;
    hgc::Staging up(g->stream);
    for (int l = 0; l < L; ++l) {
        HIP_TRY(tmp.alloc(&d_rp[size_t(l)], n1));
        HIP_TRY(tmp.alloc(&d_col[size_t(l)], m[size_t(l)]));
    }
    if (offset_bits == 32) HIP_TRY(tmp.alloc(&d_rp32, n1));
    HIP_TRY(tmp.alloc(&d_rp_arr, size_t(L)));
    HIP_TRY(tmp.alloc(&d_col_arr, size_t(L)));
    HIP_TRY(tmp.alloc(&tot, n1));
    HIP_TRY(tmp.alloc(&tot_s, n1));
    HIP_TRY(tmp.alloc(&chunk_cnt, n1));
    HIP_TRY(tmp.alloc(&coff, n1));
    HIP_TRY(tmp.alloc(&ids, size_t(n)));
    HIP_TRY(tmp.alloc(&perm, size_t(n)));
    HIP_TRY(tmp.alloc(&nzc, size_t(n)));
    HIP_TRY(tmp.alloc(&prep, size_t(kPrepWords)));
    HIP_TRY(dalloc(&g->off, n1));
    HIP_TRY(dalloc(&g->inv, size_t(n)));
    HIP_TRY(dalloc(&g->z, size_t(n)));

    for (int l = 0; l < L; ++l) {
        if (offset_bits == 64) {
            HIP_TRY(up.h2d(d_rp[size_t(l)], layers[l].row_ptr, sizeof(int64_t) * n1));
        } else {  // staged as int32, widened on the device (the copy and kernel are ordered on the stream)
            HIP_TRY(up.h2d(d_rp32, layers[l].row_ptr, sizeof(int32_t) * n1));
            hipLaunchKernelGGL(k_widen_offsets, dim3(std::max(grid_n, 1u)), dim3(kBlock), 0, g->stream, d_rp32,
                               uint64_t(n1), d_rp[size_t(l)]);
        }
        if (m[size_t(l)]) HIP_TRY(up.h2d(d_col[size_t(l)], layers[l].col, sizeof(uint32_t) * m[size_t(l)]));
    }
    HIP_TRY(up.h2d(d_rp_arr, d_rp.data(), sizeof(int64_t*) * L));
    HIP_TRY(up.h2d(d_col_arr, d_col.data(), sizeof(uint32_t*) * L));
    HIP_TRY(hipMemsetAsync(prep, 0, sizeof(unsigned long long) * kPrepWords, g->stream));
    hipLaunchKernelGGL(k_prep_vertex, dim3(std::max(grid_n, 1u)), dim3(kBlock), 0, g->stream, d_rp_arr, g->n, g->L,
                       tot, nzc, ids, prep);
    HIP_TRY(hipGetLastError());
    {
        size_t tb = 0;
        HIP_TRY(sort_by_total(nullptr, tb, tot, tot_s, ids, perm, g->n, g->stream));
        uint8_t* t = nullptr;
        HIP_TRY(tmp.alloc(&t, std::max<size_t>(tb, 1)));
        HIP_TRY(sort_by_total(t, tb, tot, tot_s, ids, perm, g->n, g->stream));
    }
    // the prefix sums run over n + 1 items: the extra last item is 0
    HIP_TRY(hipMemsetAsync(tot_s + n, 0, sizeof(uint64_t), g->stream));
    HIP_TRY(hipMemsetAsync(chunk_cnt + n, 0, sizeof(uint64_t), g->stream));
    hipLaunchKernelGGL((k_prep_rank<uint64_t(kChunk), uint64_t(kMinDefer), kPullHeavy>), dim3(std::max(grid_n, 1u)),
                       dim3(kBlock), 0, g->stream, tot_s, perm, nzc, g->n, g->inv, g->z, chunk_cnt, prep);
    HIP_TRY(hipGetLastError());
    {
        uint64_t* off_u = reinterpret_cast<uint64_t*>(g->off);
        size_t tb1 = 0, tb2 = 0;
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb1, tot_s, off_u, int(n1), g->stream));
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, chunk_cnt, coff, int(n1), g->stream));
        uint8_t* t = nullptr;
        HIP_TRY(tmp.alloc(&t, std::max<size_t>(std::max(tb1, tb2), 1)));
        size_t tb = std::max(tb1, tb2);
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(t, tb, tot_s, off_u, int(n1), g->stream));
        tb = std::max(tb1, tb2);
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(t, tb, chunk_cnt, coff, int(n1), g->stream));
    }
    unsigned long long h_prep[kPrepWords];
    uint64_t h_end[2];
    HIP_TRY(up.d2h(h_prep, prep, sizeof h_prep));
    HIP_TRY(up.d2h(&h_end[0], g->off + n, sizeof(uint64_t)));
    HIP_TRY(up.d2h(&h_end[1], coff + n, sizeof(uint64_t)));
    if (h_prep[kPrepBad] & kBadRowPtr) return fail(HG_ERR_ARG, "row_ptr must be non-decreasing");
    if (h_end[0] != D) return fail(HG_ERR_DEVICE, "set-up: merged offsets do not sum to the entry count");
    const uint64_t maxdeg = h_prep[kPrepMaxDeg];
    g->nchunks = h_end[1];
    g->pcap = h_prep[kPrepDefer] + 1;
    g->max_nz = uint32_t(h_prep[kPrepMaxNz]);
    g->H = uint32_t(h_prep[kPrepHeavy]);
    g->pull = uint32_t(L) <= kPullMaxL && env_int("HG_PULL", 1) == 1 ? 1u : 0u;
    g->kcap = uint32_t(maxdeg + 2);
    // Row compaction between levels (k_cmp_*), off by default: on the
    // 100M-vertex workload it cuts the steps from 3229 to 2301 and their time
    // by ~40 ms, but its passes over all n vertices cost more (598-636 vs
    // 545 ms per decomposition).  HG_COMPACT=1 turns it on for graphs of 2^28
    // entries or more, 2 for every graph (tests).
    {
        const char* cv = std::getenv("HG_COMPACT");
        const int c = cv && *cv ? std::atoi(cv) : 0;
        g->compact_min = c == 1 ? (1ull << 28) : c == 2 ? 1 : 0;
        g->compact_ratio = uint64_t(std::max(2, env_int("HG_CMP_RATIO", 2)));
    }
    // Level-kernel grid: 3 blocks per CU (the cascade's LDS allows 3; a sweep
    // on the 1M- and 100M-vertex workloads picked it over 2 and 4); HG_BLOCKS
    // overrides.
    const uint32_t want_g = uint32_t(env_int("HG_BLOCKS", 3 * g->cus));
    g->G = std::max(1u, std::min<uint32_t>(want_g, uint32_t(n)));
    // list slice per block: its runs of kRun ids, dealt round robin
    g->cap = uint32_t(((uint64_t(n) + kRun - 1) / kRun + g->G - 1) / g->G * kRun);

    const size_t nL = size_t(n) * size_t(L);
    int64_t bytes = 0;
    auto acct = [&](size_t b) { bytes += int64_t(hgc::pool_block_bytes(b)); };  // as the cache hands blocks out
    acct(8 * n1);  // off, inv, z
    acct(4 * size_t(n));
    acct(size_t(n));
    HIP_TRY(dalloc(&g->adj, g->D)); acct(4 * g->D);
    HIP_TRY(dalloc(&g->adjz, g->D + 16)); acct(g->D);  // the pull reads whole 16-byte words
    HIP_TRY(dalloc(&g->deg0, nL)); acct(4 * nL);
    const size_t aw = (size_t(n) + 31) / 32;
    HIP_TRY(dalloc(&g->alive, aw * L)); acct(4 * aw * L);
    HIP_TRY(dalloc(&g->out, nL)); acct(4 * nL);
    HIP_TRY(dalloc(&g->chunks, g->nchunks)); acct(8 * g->nchunks);
    HIP_TRY(dalloc(&g->deg, nL * L)); acct(4 * nL * L);
    HIP_TRY(dalloc(&g->cnt, nL)); acct(4 * nL);
    HIP_TRY(dalloc(&g->idx, nL)); acct(4 * nL);
    const size_t gcap = size_t(g->G) * g->cap;
    HIP_TRY(dalloc(&g->lists, gcap * L)); acct(4 * gcap * L);
    HIP_TRY(dalloc(&g->rings, gcap * L)); acct(4 * gcap * L);
    HIP_TRY(dalloc(&g->blk, size_t(2) * g->G * L)); acct(8 * size_t(g->G) * L);
    HIP_TRY(dalloc(&g->pl, size_t(2) * g->pcap * L)); acct(16 * size_t(g->pcap) * L);
    HIP_TRY(dalloc(&g->d_ihist, size_t(g->kcap) * L)); acct(8 * size_t(g->kcap) * L);
    HIP_TRY(dalloc(&g->d_insts, size_t(L)));
    HIP_TRY(dalloc(&g->d_states, size_t(L)));
    HIP_TRY(dalloc(&g->d_ctl, 1));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&g->h_ctl), sizeof(Control), hipHostMallocDefault));
    HIP_TRY(dalloc(&g->d_density, size_t(g->kcap)));
    HIP_TRY(dalloc(&g->d_sel, size_t(g->kcap)));
    HIP_TRY(dalloc(&g->d_cell_len, size_t(g->kcap)));
    HIP_TRY(dalloc(&g->d_pow, size_t(L)));
    for (auto& e : g->ev) HIP_TRY(hipEventCreate(&e));
    g->device_bytes = bytes;

    // deg0 / adj / chunks in internal ids
    const uint32_t grid = uint32_t(std::min<uint64_t>((nL + kBlock - 1) / kBlock, uint64_t(g->cus) * 16));
    hipLaunchKernelGGL(k_build_deg0, dim3(std::max(grid, 1u)), dim3(kBlock), 0, g->stream, d_rp_arr, perm, g->n, g->L,
                       g->deg0);
    const uint32_t grid2 = uint32_t(std::min<uint64_t>((uint64_t(n) * kWave + kBlock - 1) / kBlock,
                                                       uint64_t(g->cus) * 16));
    hipLaunchKernelGGL(k_build_adj, dim3(std::max(grid2, 1u)), dim3(kBlock), 0, g->stream, d_rp_arr, d_col_arr, g->deg0,
                       g->off, perm, g->inv, nzc, g->n, g->L, g->adj, g->adjz, prep);
    hipLaunchKernelGGL(k_build_chunks, dim3(std::max(grid_n, 1u)), dim3(kBlock), 0, g->stream, g->off, coff, g->n,
                       g->chunks);
    HIP_TRY(hipGetLastError());
    HIP_TRY(up.d2h(h_prep, prep, sizeof h_prep));  // waits for the build kernels
    if (h_prep[kPrepBad] & kBadCol) return fail(HG_ERR_ARG, "a neighbour id is >= n");
    if (env_int("HG_VERIFY_UPLOAD", 0) == 1) {  // diagnostic: read every layer back
        for (int l = 0; l < L; ++l) {
            const size_t ml = size_t(m[size_t(l)]);
            std::vector<int64_t> brp(n1);
            std::vector<uint32_t> bcol(ml);
            HIP_TRY(up.d2h(brp.data(), d_rp[size_t(l)], sizeof(int64_t) * n1));
            if (ml) HIP_TRY(up.d2h(bcol.data(), d_col[size_t(l)], sizeof(uint32_t) * ml));
            bool same = ml == 0 || std::memcmp(bcol.data(), layers[l].col, sizeof(uint32_t) * ml) == 0;
            for (size_t v = 0; v < n1 && same; ++v)
                same = brp[v] == (offset_bits == 64 ? static_cast<const int64_t*>(layers[l].row_ptr)[v]
                                                    : int64_t(static_cast<const int32_t*>(layers[l].row_ptr)[v]));
            if (!same) return fail(HG_ERR_HIP, "upload verify: layer " + std::to_string(l) + " differs on the device");
        }
    }
    const uint32_t lmagic = g->L == 1 ? 0xFFFFFFFFu : uint32_t((1ull << 32) / g->L);  // div_L needs magic < 2^32
    if (hgc::hb_enabled(g->D, g->L)) {
        // core-index bounds per entry, from the per-vertex ones (freed on return)
        uint8_t* hv = nullptr;
        HIP_TRY(tmp.alloc(&hv, nL));
        HIP_TRY(dalloc(&g->adjh, size_t(L) * g->D));
        g->device_bytes += int64_t(hgc::hb_bytes(g->D, g->L));
        const uint32_t gv = uint32_t(std::min<uint64_t>((uint64_t(n) * kWave + kBlock - 1) / kBlock,
                                                        uint64_t(g->cus) * 16));
        hipLaunchKernelGGL(k_build_vbound, dim3(std::max(gv, 1u)), dim3(kBlock), 0, g->stream, g->off, g->deg0,
                           g->adjz, g->n, g->L, hv);
        const uint32_t ge = uint32_t(std::min<uint64_t>((g->D + kBlock - 1) / kBlock, uint64_t(g->cus) * 16));
        hipLaunchKernelGGL(k_build_adjh, dim3(std::max(ge, 1u)), dim3(kBlock), 0, g->stream, g->adj, g->D, g->L,
                           lmagic, hv, g->adjh);
        HIP_TRY(hipGetLastError());
    }
    if (hgc::bin_enabled(g->D)) {
        int rc = bin_setup(g, tmp, up);
        if (rc) return rc;
        g->device_bytes += int64_t(hgc::bin_bytes(g->D, g->L));
    }
    // the temporaries are freed on return: nothing queued reads them any more
    HIP_TRY(hipStreamSynchronize(g->stream));

    g->insts.resize(size_t(L));
    for (uint32_t i = 0; i < g->L; ++i) {
        Inst& in = g->insts[i];
        in.off = g->off;
        in.adj = g->adj;
        in.adjz = g->adjz;
        in.adjh = g->adjh ? g->adjh + size_t(i) * g->D : nullptr;
        in.alive = g->alive + aw * i;
        in.deg = g->deg + size_t(i) * nL;
        in.cnt = g->cnt + size_t(i) * n;
        in.idx = g->idx + size_t(i) * n;
        in.list = g->lists + size_t(i) * gcap;
        in.ring = g->rings + size_t(i) * gcap;
        in.blk = g->blk + size_t(2) * g->G * i;
        in.pl0 = g->pl + size_t(2 * i) * g->pcap;
        in.pl1 = g->pl + size_t(2 * i + 1) * g->pcap;
        in.plcap = g->pcap;
        in.n = g->n;
        in.L = g->L;
        in.lmd = i + 1;
        in.cap = g->cap;
        in.lmagic = lmagic;
        in.flags = env_int("HG_AGG", 1) == 2 ? 0u : kFlagAggregate;  // HG_AGG=2 disables combining
        in.heavy = uint32_t(env_int("HG_HEAVY", int(kHeavy)));
        in.tail_heavy = uint32_t(env_int("HG_TAIL_HEAVY", int(in.heavy)));
        in.tail_chunks = uint32_t(env_int("HG_TAIL_CHUNKS", int(g->G)));
        // deferred chunk size; the lists are sized for chunks of kMinDefer
        in.defer = std::min<uint32_t>(std::max<uint32_t>(uint32_t(env_int("HG_DEFER", int(kDefer))), kMinDefer),
                                      1u << 23);
    }
    return HG_OK;
}

int sync_ctl(hg_graph* g) {
    HIP_TRY(hipMemcpyAsync(g->h_ctl, g->d_ctl, sizeof(Control), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return HG_OK;
}

const char* err_name(int32_t e) {
    switch (e) {
        case kErrDefer: return "deferred chunk list overflow";
        case kErrCompact: return "row compaction: copied rows do not match the degrees";
        case kErrBin: return "binned apply: a record fell outside its bin";
        default: return "unknown device error";
    }
}

// Event `i` of the phase-timing pool, created on first use.
int phase_event(hg_graph* g, size_t i) {
    while (g->pev.size() <= i) {
        hipEvent_t e = nullptr;
        HIP_TRY(hipEventCreate(&e));
        g->pev.push_back(e);
    }
    return HG_OK;
}

// Launches step s: the scan (level-0 init at s = 0), which also decides
// the step, then the cascade.  Kernels of instances that are done return at
// once.  With phase timing, events 2s and 2s+1 go before the scan and
// before the cascade.
int launch_step(hg_graph* g, uint32_t s) {
    if (g->phase_timing) {
        int rc = phase_event(g, 2 * size_t(s) + 1);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(g->pev[2 * size_t(s)], g->stream));
    }
    const uint32_t L = g->L;
    if (s == 0) {
        // level-0 counters per thread: 8 when the graph has at most 8 layers
        // (fewer registers, so three blocks fit a CU), else 16
        const bool narrow = L <= 8;
        if (g->pull && g->H > 0) {
            const uint32_t gh = uint32_t(std::min<uint64_t>((uint64_t(g->H) * kWave + kBlock - 1) / kBlock,
                                                            uint64_t(g->cus) * 8));
            if (narrow)
                hipLaunchKernelGGL(k_pull_heavy<8>, dim3(gh), dim3(kBlock), 0, g->stream, g->d_insts, L, g->deg0, g->z,
                                   g->H, g->L);
            else
                hipLaunchKernelGGL(k_pull_heavy<16>, dim3(gh), dim3(kBlock), 0, g->stream, g->d_insts, L, g->deg0,
                                   g->z, g->H, g->L);
        }
        if (narrow)
            hipLaunchKernelGGL(k_init<8>, dim3(g->G), dim3(kBlock), 0, g->stream, g->d_insts, g->d_states, g->d_ctl, L,
                               g->deg0, g->z, g->n, g->L, g->cap, g->H, g->pull);
        else
            hipLaunchKernelGGL(k_init<16>, dim3(g->G), dim3(kBlock), 0, g->stream, g->d_insts, g->d_states, g->d_ctl,
                               L, g->deg0, g->z, g->n, g->L, g->cap, g->H, g->pull);
    }
    else
        hipLaunchKernelGGL(k_scan, dim3(g->G), dim3(kBlock), 0, g->stream, g->d_insts, g->d_states, g->d_ctl, L, s);
    if (g->phase_timing) HIP_TRY(hipEventRecord(g->pev[2 * size_t(s) + 1], g->stream));
    // diagnostic (HG_BIN_LOG=1): every step's kernels timed one by one and
    // the binned instances' frontier entries, records and removals printed
    static const bool blog = env_int("HG_BIN_LOG", 0) == 1;
    double tk[4] = {0, 0, 0, 0}, tsum[3] = {0, 0, 0};
    auto tick = [&](int slot, std::chrono::steady_clock::time_point t0) -> int {
        HIP_TRY(hipStreamSynchronize(g->stream));
        tk[slot] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return HG_OK;
    };
    std::vector<InstState> before;
    if (blog) {
        HIP_TRY(hipStreamSynchronize(g->stream));
        before.resize(L);
        HIP_TRY(hipMemcpy(before.data(), g->d_states, sizeof(InstState) * L, hipMemcpyDeviceToHost));
    }
    if (g->bin_on && s <= g->bin_steps && (s >= 1 || g->bin_level0)) {
        // dense removal rounds: binned apply (the kernels return at once for
        // instances that do not bin this step)
        const uint32_t g2 = 2u * uint32_t(g->cus);
        for (uint32_t r = 0; r < g->bin_rounds; ++r) {  // removal rounds 0, 1, ... of the level
            double t[3] = {0, 0, 0};
            auto t0 = std::chrono::steady_clock::now();
            const uint32_t gx = g->bin_xg * uint32_t(g->cus);
            if (g->bin_eu == 8)
                hipLaunchKernelGGL(k_bin_expand<8>, dim3(gx), dim3(kBlock), 0, g->stream, g->d_insts, g->d_states,
                                   g->d_ctl, L, s, r, g->bgeo);
            else
                hipLaunchKernelGGL(k_bin_expand<4>, dim3(gx), dim3(kBlock), 0, g->stream, g->d_insts, g->d_states,
                                   g->d_ctl, L, s, r, g->bgeo);
            if (blog && tick(0, t0)) return HG_ERR_HIP;
            t[0] = tk[0];
            t0 = std::chrono::steady_clock::now();
            if (g->bin_sg == 3)
                hipLaunchKernelGGL(k_bin_scatter<kBinScatterU3>, dim3(3u * uint32_t(g->cus)), dim3(kBlock), 0,
                                   g->stream, g->d_insts, g->d_states, g->d_ctl, L, s, r, g->bgeo);
            else
                hipLaunchKernelGGL(k_bin_scatter<kBinScatterU2>, dim3(g2), dim3(kBlock), 0, g->stream, g->d_insts,
                                   g->d_states, g->d_ctl, L, s, r, g->bgeo);
            if (blog && tick(1, t0)) return HG_ERR_HIP;
            t[1] = tk[1];
            t0 = std::chrono::steady_clock::now();
            const uint32_t ga = g->bin_ag * uint32_t(g->cus);
            hipLaunchKernelGGL(k_bin_group<kBinFineU>, dim3(ga), dim3(kBlock), 0, g->stream, g->d_states, g->d_ctl, L,
                               r, g->bgeo);
            if (g->bin_ag == 4)
                hipLaunchKernelGGL(k_bin_fine<kBinDeltaWords4>, dim3(ga), dim3(kBlock), 0, g->stream, g->d_insts,
                                   g->d_states, g->d_ctl, L, s, r, g->bgeo);
            else if (g->bin_ag == 3)
                hipLaunchKernelGGL(k_bin_fine<kBinDeltaWords3>, dim3(ga), dim3(kBlock), 0, g->stream, g->d_insts,
                                   g->d_states, g->d_ctl, L, s, r, g->bgeo);
            else
                hipLaunchKernelGGL(k_bin_fine<kBinDeltaWords2>, dim3(ga), dim3(kBlock), 0, g->stream, g->d_insts,
                                   g->d_states, g->d_ctl, L, s, r, g->bgeo);
            if (blog && tick(2, t0)) return HG_ERR_HIP;
            t[2] = tk[2];
            for (int q = 0; q < 3; ++q) tsum[q] += t[q];
        }
        for (int q = 0; q < 3; ++q) tk[q] = tsum[q];
    }
    auto t3 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(k_cascade, dim3(g->G), dim3(kBlock), 0, g->stream, g->d_insts, g->d_states, g->d_ctl, L, s);
    if (blog) {
        if (tick(3, t3)) return HG_ERR_HIP;
        std::vector<InstState> after(L);
        HIP_TRY(hipMemcpy(after.data(), g->d_states, sizeof(InstState) * L, hipMemcpyDeviceToHost));
        const uint32_t p = s & 1u;
        std::string line;
        for (uint32_t i = 0; i < L; ++i) {
            const InstState& a = after[i];
            if (a.dphase[p] == kPhaseDone) continue;
            char buf[256];
            std::snprintf(buf, sizeof buf, " | l%u %s%d k%d f0ent %llu rec %llu lastch %llu alive %llu", i + 1,
                          a.dphase[p] == kPhaseScan ? (a.bmode[p] ? "B" : "S") : "C", a.bmode[p] ? a.bround[p] : -1,
                          a.dk[p], a.pfent[p][0],
                          a.brec[0] - before[i].brec[0], a.bmode[p] ? a.pbr[a.bround[p] & 1][0] : 0ull, a.alive[0]);
            line += buf;
        }
        Control c{};
        HIP_TRY(hipMemcpy(&c, g->d_ctl, sizeof(Control), hipMemcpyDeviceToHost));
        std::fprintf(stderr, "[hg] step %u expand %.3f scatter %.3f apply %.3f cascade %.3f scatter-batches %llu%s\n", s,
                     tk[0], tk[1], tk[2], tk[3], c.dbg[0], line.c_str());
    }
    return HG_OK;
}

// Grows *p to hold at least `need` items (contents are not kept).
template <typename T>
int ensure_cap(hg_graph* g, T** p, uint64_t* cap, uint64_t need) {
    if (*p && *cap >= need) return HG_OK;
    if (*p) {
        hgc::dfree(*p);
        g->cmp_bytes -= int64_t(*cap * sizeof(T));
        *p = nullptr;
        *cap = 0;
    }
    const uint64_t want = need + need / 8 + 1024;
    if (dalloc(p, size_t(want)) != hipSuccess) return fail(HG_ERR_NOMEM, "row compaction: device allocation failed");
    *cap = want;
    g->cmp_bytes += int64_t(want * sizeof(T));
    return HG_OK;
}

uint32_t cmp_grid(const hg_graph* g, uint64_t items_per_thread_space) {
    return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>((items_per_thread_space + kBlock - 1) / kBlock,
                                                             uint64_t(g->cus) * 16)));
}

// Sum over alive vertices of instance i of sum_l deg[v][l] (= its alive
// entries at a level boundary); the arrays are written too when `write`.
int cmp_count(hg_graph* g, const Inst& in, bool write, uint64_t* sum) {
    HIP_TRY(hipMemsetAsync(g->c_words, 0, sizeof(unsigned long long) * 4, g->stream));
    const uint32_t grid = cmp_grid(g, uint64_t(g->n) + 1);
    if (write)
        hipLaunchKernelGGL(k_cmp_count<true>, dim3(grid), dim3(kBlock), 0, g->stream, in, g->c_len, g->c_nch,
                           g->c_words);
    else
        hipLaunchKernelGGL(k_cmp_count<false>, dim3(grid), dim3(kBlock), 0, g->stream, in, g->c_len, g->c_nch,
                           g->c_words);
    HIP_TRY(hipGetLastError());
    unsigned long long w = 0;
    HIP_TRY(hipMemcpyAsync(&w, g->c_words, sizeof w, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    *sum = w;
    return HG_OK;
}

// Rebuilds instance i's rows with only alive targets into its other buffer
// (see k_cmp_*); `alive_entries` is the new total.
int compact_rows(hg_graph* g, uint32_t i, uint64_t alive_entries) {
    const size_t n1 = size_t(g->n) + 1;
    hg_graph::CmpBuf& cb = g->cmp[i];
    Inst& in = g->cur[i];
    uint64_t sum = 0;
    int rc = cmp_count(g, in, true, &sum);
    if (rc) return rc;
    if (sum != alive_entries) return fail(HG_ERR_DEVICE, "row compaction: alive entry count changed");
    const int dst = in.off == cb.off[0] ? 1 : 0;
    if (!cb.off[dst]) {
        if (dalloc(&cb.off[dst], n1) != hipSuccess) return fail(HG_ERR_NOMEM, "row compaction: device allocation failed");
        g->cmp_bytes += int64_t(8 * n1);
    }
    if ((rc = ensure_cap(g, &cb.adj[dst], &cb.adj_cap[dst], std::max<uint64_t>(sum, 1)))) return rc;
    {
        size_t tb = 0;
        uint64_t* off_u = reinterpret_cast<uint64_t*>(cb.off[dst]);
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, g->c_len, off_u, int(n1), g->stream));
        if (tb > g->c_temp_bytes) {
            if (g->c_temp) hgc::dfree(g->c_temp);
            g->c_temp = nullptr;
            if (dalloc(&g->c_temp, tb) != hipSuccess) return fail(HG_ERR_NOMEM, "row compaction: device allocation failed");
            g->c_temp_bytes = tb;
        }
        size_t t1 = g->c_temp_bytes;
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(g->c_temp, t1, g->c_len, off_u, int(n1), g->stream));
        size_t t2 = g->c_temp_bytes;
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(g->c_temp, t2, g->c_nch, g->c_item_off, int(n1), g->stream));
    }
    uint64_t nitems = 0;
    HIP_TRY(hipMemcpyAsync(&nitems, g->c_item_off + g->n, sizeof nitems, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    if ((rc = ensure_cap(g, &g->c_items, &g->c_items_cap, std::max<uint64_t>(nitems, 1)))) return rc;
    hipLaunchKernelGGL(k_cmp_items, dim3(cmp_grid(g, g->n)), dim3(kBlock), 0, g->stream, in, cb.off[dst],
                       g->c_item_off, g->c_items, g->c_len);
    const uint32_t fill_grid = uint32_t(std::max<uint64_t>(
        1, std::min<uint64_t>((nitems * kWave + kBlock - 1) / kBlock, uint64_t(g->cus) * 16)));
    hipLaunchKernelGGL(k_cmp_fill, dim3(fill_grid), dim3(kBlock), 0, g->stream, in, g->c_items, nitems, cb.off[dst],
                       g->c_len, cb.adj[dst], g->d_ctl);
    hipLaunchKernelGGL(k_cmp_check, dim3(cmp_grid(g, g->n)), dim3(kBlock), 0, g->stream, in, cb.off[dst], g->c_len,
                       g->d_ctl);
    HIP_TRY(hipGetLastError());
    in.off = cb.off[dst];
    in.adj = cb.adj[dst];
    in.flags |= kFlagCompact;
    HIP_TRY(hipMemcpyAsync(g->d_insts + i, &in, sizeof(Inst), hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));  // `in` may change before the next compaction
    ++g->compactions;
    return HG_OK;
}

// At a host sync before step s: compacts the rows of every instance at a
// level boundary (nothing deferred to step s) whose alive entries have
// fallen to half of its rows' entries at the last compaction (the graph's
// entries at first).  The alive entries are counted only after the alive
// vertices have fallen by a quarter since the last count.
struct CmpTrack {
    uint64_t next_check;  // count again once the alive vertices are at most this
    uint64_t rows;        // entries of the current rows
};

int maybe_compact(hg_graph* g, uint32_t s, std::vector<CmpTrack>& tr) {
    const uint32_t L = g->L;
    std::vector<InstState>& st = g->h_states;
    HIP_TRY(hipMemcpyAsync(st.data(), g->d_states, sizeof(InstState) * L, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    for (uint32_t i = 0; i < L; ++i) {
        const InstState& x = st[i];
        if (x.dphase[(s + 1u) & 1u] == kPhaseDone) continue;  // the last step's decision
        if (x.pdefer[s & 1u][0] != 0) continue;               // mid-level
        const uint64_t alive = x.alive[0];
        if (alive == 0 || alive > tr[i].next_check) continue;
        tr[i].next_check = alive - alive / 4;
        uint64_t a = 0;
        const bool log = env_int("HG_CMP_LOG", 0) == 1;  // diagnostic
        const auto t0 = std::chrono::steady_clock::now();
        int rc = cmp_count(g, g->cur[i], false, &a);
        if (rc) return rc;
        if (log)
            std::fprintf(stderr, "[hg] step %u lmd %u alive %llu entries %llu rows %llu count %.3f ms\n", s, i + 1,
                         (unsigned long long)alive, (unsigned long long)a, (unsigned long long)tr[i].rows,
                         std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        if (g->compact_ratio * a > tr[i].rows) continue;
        const auto t1 = std::chrono::steady_clock::now();
        if ((rc = compact_rows(g, i, a))) return rc;
        if (log)
            std::fprintf(stderr, "[hg]   compacted lmd %u in %.3f ms\n", i + 1,
                         std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count());
        tr[i].rows = a;
    }
    return HG_OK;
}

int decompose_impl(hg_graph* g) {
    HIP_TRY(hipSetDevice(g->device));
    const uint32_t L = g->L;
    g->decomposed = false;
    g->densest_done = false;
    std::vector<InstState>& st = g->h_states;
    st.assign(L, InstState{});
    std::memset(st.data(), 0, sizeof(InstState) * L);
    for (uint32_t i = 0; i < L; ++i) {
        // step 0 decides from the slots of parity 1: level -1 + 1 = 0
        st[i].alive[0] = g->n;
        st[i].dphase[1] = kPhaseScan;
        st[i].dk[1] = -1;
        if (i + 1 > g->max_nz) {
            // No vertex has edges in lmd layers: S(1, lmd) is empty, every
            // index is 0 (k_init writes the zeros).
            st[i].dphase[1] = kPhaseDone;
            st[i].alive[0] = 0;
        }
    }
    Control& ctl = g->h_ctl_init;
    std::memset(&ctl, 0, sizeof ctl);
    ctl.active = int32_t(L);
    hgc::Staging io(g->stream);
    HIP_TRY(io.h2d(g->d_insts, g->insts.data(), sizeof(Inst) * L));
    g->cur = g->insts;  // the rows every instance starts from
    g->compactions = 0;
    const bool compact = g->compact_min != 0 && g->D >= g->compact_min;
    std::vector<CmpTrack> track(L, CmpTrack{uint64_t(g->n), g->D});
    if (compact && !g->c_len) {
        const size_t n1 = size_t(g->n) + 1;
        if (g->cmp.empty()) g->cmp.resize(L);
        if (dalloc(&g->c_len, n1) != hipSuccess || dalloc(&g->c_nch, n1) != hipSuccess ||
            dalloc(&g->c_item_off, n1) != hipSuccess || dalloc(&g->c_words, 4) != hipSuccess)
            return fail(HG_ERR_NOMEM, "row compaction: device allocation failed");
        g->cmp_bytes += int64_t(24 * n1 + 32);
    }
    HIP_TRY(io.h2d(g->d_states, st.data(), sizeof(InstState) * L));
    HIP_TRY(io.h2d(g->d_ctl, &ctl, sizeof(Control)));
    if (g->bin_on) {  // the per-bin tables start at zero (an earlier call may have stopped on an error)
        const size_t tab = size_t(L) * g->bgeo.cmax;
        HIP_TRY(hipMemsetAsync(g->bin_bh, 0, sizeof(uint32_t) * tab, g->stream));
        HIP_TRY(hipMemsetAsync(g->bin_bc, 0, sizeof(uint32_t) * tab, g->stream));
    }

    HIP_TRY(hipEventRecord(g->ev[0], g->stream));
    uint32_t s = 0;
    int lrc = launch_step(g, s++);
    if (lrc) return lrc;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(g->ev[1], g->stream));
    // Steps run in batches; the host only reads the active count between
    // batches (a finished instance makes its kernels return at once).
    static const int batch_max = env_int("HG_BATCH", 32);
    uint32_t batch = 4;
    const uint64_t step_bound = 4ull * (uint64_t(g->kcap) + 8) + 2ull * g->n + 64;
    for (;;) {
        for (uint32_t b = 0; b < batch; ++b) {
            lrc = launch_step(g, s++);
            if (lrc) return lrc;
        }
        HIP_TRY(hipGetLastError());
        int rc = sync_ctl(g);
        if (rc) return rc;
        if (g->h_ctl->error) return fail(HG_ERR_DEVICE, err_name(g->h_ctl->error));
        if (g->h_ctl->active == 0) break;
        if (compact) {
            rc = maybe_compact(g, s, track);
            if (rc) return rc;
        }
        if (s > step_bound) return fail(HG_ERR_DEVICE, "step loop did not terminate");
        batch = std::min<uint32_t>(batch * 2, uint32_t(batch_max));
    }
    if (g->phase_timing) {
        lrc = phase_event(g, 2 * size_t(s));
        if (lrc) return lrc;
        HIP_TRY(hipEventRecord(g->pev[2 * size_t(s)], g->stream));
    }
    // Cell lengths and kmax from the histogram of core indices.
    HIP_TRY(hipMemsetAsync(g->d_ihist, 0, sizeof(uint64_t) * size_t(g->kcap) * L, g->stream));
    const uint32_t hx = uint32_t(std::min<uint64_t>((uint64_t(g->n) + kBlock - 1) / kBlock, uint64_t(g->cus) * 2));
    const uint32_t ux = uint32_t(std::min<uint64_t>((uint64_t(g->n) * L + kBlock - 1) / kBlock, uint64_t(g->cus) * 16));
    hipLaunchKernelGGL(k_unpermute, dim3(std::max(ux, 1u)), dim3(kBlock), 0, g->stream, g->d_insts, g->inv, g->n, L,
                       g->out);
    hipLaunchKernelGGL(k_idx_hist, dim3(std::max(hx, 1u), L), dim3(kBlock), 0, g->stream, g->out, g->n, L, g->d_ihist,
                       g->kcap);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(g->ev[2], g->stream));
    HIP_TRY(hipEventSynchronize(g->ev[2]));
    float ms_all = 0, ms_init = 0;
    HIP_TRY(hipEventElapsedTime(&ms_all, g->ev[0], g->ev[2]));
    HIP_TRY(hipEventElapsedTime(&ms_init, g->ev[0], g->ev[1]));
    g->timing.decompose_ms = ms_all;
    g->timing.init_ms = ms_init;
    g->timing.scan_ms = -1;
    g->timing.cascade_ms = -1;
    if (g->phase_timing) {
        // step 0 (level-0 init and its cascade) is init_ms, as without
        // phase timing; scans and cascades of steps 1..s-1 are summed
        double scan = 0, cas = 0;
        const bool plog = env_int("HG_PHASE_LOG", 0) == 1;  // diagnostic: every step's two phases
        for (uint32_t t = 0; t < s; ++t) {
            float a = 0, b = 0;
            HIP_TRY(hipEventElapsedTime(&a, g->pev[2 * size_t(t)], g->pev[2 * size_t(t) + 1]));
            HIP_TRY(hipEventElapsedTime(&b, g->pev[2 * size_t(t) + 1], g->pev[2 * size_t(t) + 2]));
            if (t == 0) {
                ms_init = a + b;
            } else {
                scan += a;
                cas += b;
            }
            if (plog) std::fprintf(stderr, "[hg] step %u scan %.4f cascade %.4f\n", t, a, b);
        }
        g->timing.init_ms = ms_init;
        g->timing.scan_ms = scan;
        g->timing.cascade_ms = cas;
    }
    g->timing.levels = int64_t(s);
    g->timing.binned = g->h_ctl->binned;
    {
        // steps s0..min(bin_steps, s-1) launched the binned kernels too
        const int64_t s0 = g->bin_level0 ? 0 : 1, s1 = std::min<int64_t>(g->bin_steps, int64_t(s) - 1);
        const int64_t nb = g->bin_on && s1 >= s0 ? s1 - s0 + 1 : 0;
        g->timing.launches = int64_t(s) * 2 + 3 + 3 * int64_t(g->bin_rounds) * nb;
    }

    g->h_count.assign(size_t(g->kcap) * L, 0);
    HIP_TRY(io.d2h(g->h_count.data(), g->d_ihist, sizeof(int64_t) * g->h_count.size()));
    g->h_kmax.assign(L, 0);
    int64_t total = 0;
    for (uint32_t i = 0; i < L; ++i) {
        for (uint32_t k = 0; k < g->kcap; ++k) {
            const int64_t c = g->h_count[size_t(i) * g->kcap + k];
            total += c;
            if (c > 0) g->h_kmax[i] = int32_t(k);
        }
    }
    if (total != int64_t(g->n) * L)
        return fail(HG_ERR_DEVICE, "core index histogram covers " + std::to_string(total) + " of " +
                                       std::to_string(int64_t(g->n) * L) + " entries");
    g->decomposed = true;
    return HG_OK;
}

std::vector<int64_t> cell_lengths(const hg_graph* g, uint32_t i) {
    // |S(k)| = #{v : index >= k}, k = 0..kmax
    const int32_t km = g->h_kmax[i];
    std::vector<int64_t> len(size_t(km) + 1);
    int64_t acc = 0;
    for (int32_t k = km; k >= 0; --k) {
        acc += g->h_count[size_t(i) * g->kcap + size_t(k)];
        len[size_t(k)] = acc;
    }
    return len;
}

// Device buffers of one densest pass: per-layer edge histograms [L][kc]
// (partial results), and per cell the lengths, scores and chosen layer
// counts, plus the pow(l+1, beta) table.
struct DenseBufs {
    uint64_t* hist;
    double* density;
    int32_t* sel;
    int64_t* cell_len;
    double* pow;
};

// Host <-> device copies of the densest pass: through the pinned staging
// pool, or (no_alloc) as plain copies of the caller's pageable memory, which
// allocate nothing.
int dense_h2d(hgc::Staging& io, bool no_alloc, void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (no_alloc) {
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));  // src may be reused on return
    } else {
        HIP_TRY(io.h2d(dst, src, bytes));
    }
    return HG_OK;
}

int dense_d2h(hgc::Staging& io, bool no_alloc, void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (no_alloc) {
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    } else {
        HIP_TRY(io.d2h(dst, src, bytes));
    }
    return HG_OK;
}

int densest_run(hg_graph* g, double beta, hg_wds* out, const DenseBufs& b, bool no_alloc) {
    const uint32_t L = g->L;
    std::vector<double>& pw = g->h_pow;
    pw.assign(L, 0.0);
    for (uint32_t l = 0; l < L; ++l) pw[l] = std::pow(double(l + 1), beta);
    hgc::Staging io(g->stream);
    int rc = dense_h2d(io, no_alloc, b.pow, pw.data(), sizeof(double) * L, g->stream);
    if (rc) return rc;
    g->h_density.assign(size_t(g->kcap) * L, 0.0);
    g->h_sel.assign(size_t(g->kcap) * L, 0);
    HIP_TRY(hipEventRecord(g->ev[0], g->stream));
    const uint32_t grid = uint32_t(std::min<uint64_t>((g->nchunks * kChunk / 64 + kHistBlock - 1) / kHistBlock,
                                                       uint64_t(g->cus) * 8));
    for (uint32_t i = 0; i < L; ++i) {
        const int32_t km = g->h_kmax[i];
        if (km < 1) continue;
        const uint32_t kc = uint32_t(km) + 1;
        HIP_TRY(hipMemsetAsync(b.hist, 0, sizeof(uint64_t) * size_t(L) * kc, g->stream));
        hipLaunchKernelGGL(k_edge_hist, dim3(std::max(grid, 1u)), dim3(kHistBlock), 0, g->stream, g->insts[i], g->chunks,
                           g->nchunks, b.hist, kc);
        std::vector<int64_t>& len = g->h_len;
        len = cell_lengths(g, i);
        if ((rc = dense_h2d(io, no_alloc, b.cell_len, len.data(), sizeof(int64_t) * len.size(), g->stream))) return rc;
        const uint32_t cells = uint32_t(km);
        hipLaunchKernelGGL(k_cell_score, dim3((cells + 63) / 64), dim3(64), 0, g->stream, b.hist, kc, L, uint32_t(km),
                           b.cell_len, b.pow, b.density, b.sel);
        HIP_TRY(hipGetLastError());
        if ((rc = dense_d2h(io, no_alloc, g->h_density.data() + size_t(i) * g->kcap, b.density,
                            sizeof(double) * (size_t(km) + 1), g->stream)))
            return rc;
        if ((rc = dense_d2h(io, no_alloc, g->h_sel.data() + size_t(i) * g->kcap, b.sel,
                            sizeof(int32_t) * (size_t(km) + 1), g->stream)))
            return rc;
        g->h_density[size_t(i) * g->kcap] = 0.0;
        g->h_sel[size_t(i) * g->kcap] = 0;
    }
    HIP_TRY(hipEventRecord(g->ev[1], g->stream));
    HIP_TRY(hipEventSynchronize(g->ev[1]));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, g->ev[0], g->ev[1]));
    g->timing.densest_ms = ms;
    // Argmax exactly as FCTree::WeightDenestSubgraph with main.cpp's float
    // running max: pre-order = lambda-major, k ascending; update on >=.
    hg_wds r{};
    float maxd = 0.0f;
    for (uint32_t i = 0; i < L; ++i) {
        const std::vector<int64_t> len = cell_lengths(g, i);
        for (int32_t k = 1; k <= g->h_kmax[i]; ++k) {
            const double d = g->h_density[size_t(i) * g->kcap + size_t(k)];
            if (d >= double(maxd)) {
                maxd = float(d);
                r.res_k = uint32_t(k);
                r.res_lmd = i + 1;
                r.res_len = uint32_t(len[size_t(k)]);
                r.num_select_layer = g->h_sel[size_t(i) * g->kcap + size_t(k)];
            }
        }
    }
    r.max_density = maxd;
    g->wds = r;
    g->densest_done = true;
    if (out) *out = r;
    return HG_OK;
}

int densest_impl(hg_graph* g, double beta, hg_wds* out) {
    HIP_TRY(hipSetDevice(g->device));
    if (!g->hist) {
        hipError_t e = dalloc(&g->hist, size_t(g->L) * g->kcap);
        if (e != hipSuccess) return fail(HG_ERR_NOMEM, "densest histogram allocation failed");
    }
    return densest_run(g, beta, out, DenseBufs{g->hist, g->d_density, g->d_sel, g->d_cell_len, g->d_pow}, false);
}

// ------------------------------------------------------- no-allocation densest
// hg_densest_noalloc keeps the pass's buffers in one file-scope device array
// (each GPU holds its own copy until the process ends).  In host code the
// name `scratch` is only a host-side stand-in: kernels, memsets and copies
// take the per-device address from hipGetSymbolAddress, looked up once per
// device with that device current.
static __device__ __align__(256) uint8_t scratch[HG_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_address(int device, uint8_t** out) {
    if (device < 0 || device >= kMaxDevices) return fail(HG_ERR_RANGE, "device ordinal above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(HG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = static_cast<uint8_t*>(g_scratch_ptr[device]);
    return HG_OK;
}

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

int densest_noalloc_impl(hg_graph* g, double beta, hg_wds* out) {
    HIP_TRY(hipSetDevice(g->device));
    int32_t km = 0;
    for (uint32_t i = 0; i < g->L; ++i) km = std::max(km, g->h_kmax[i]);
    const size_t kc = size_t(km) + 1;
    const size_t o_hist = 0;
    const size_t o_dens = align_up(o_hist + sizeof(uint64_t) * size_t(g->L) * kc);
    const size_t o_sel = align_up(o_dens + sizeof(double) * kc);
    const size_t o_len = align_up(o_sel + sizeof(int32_t) * kc);
    const size_t o_pow = align_up(o_len + sizeof(int64_t) * kc);
    const size_t need = o_pow + sizeof(double) * g->L;
    if (need > size_t(HG_NOALLOC_SCRATCH_BYTES))
        return fail(HG_ERR_RANGE, "hg_densest_noalloc: the pass needs " + std::to_string(need) +
                                      " bytes of scratch (8 * L * (kmax + 1) for the edge histograms plus 28 * "
                                      "(kmax + 1) + 8 * L), above HG_NOALLOC_SCRATCH_BYTES = " +
                                      std::to_string(HG_NOALLOC_SCRATCH_BYTES));
    uint8_t* base = nullptr;
    int rc = scratch_address(g->device, &base);
    if (rc) return rc;
    const DenseBufs b{reinterpret_cast<uint64_t*>(base + o_hist), reinterpret_cast<double*>(base + o_dens),
                      reinterpret_cast<int32_t*>(base + o_sel), reinterpret_cast<int64_t*>(base + o_len),
                      reinterpret_cast<double*>(base + o_pow)};
    return densest_run(g, beta, out, b, true);
}


// CoreIndex::WdsCoreIndex (CoreIndexParallel/CoreIndex.cpp:307-478) on the
// device's edge histograms.  Per cell S = S(k, lmd), k = 1..kmax, and layer
// l the reference increments a float once per directed entry inside S: the
// count is exact below 2^24 and stays at 2^24 from there (adding 1 to 2^24
// rounds back to 2^24), so n_edge = min(count, 2^24).  degree = n_edge / 2,
// sorted descending; density = (float)((degree / (float)|S|) * pow(l+1,
// (float)beta)); per lmd the record takes (density, k, lmd) on >= from
// {0, 0, 0}.  The printed answer: records [0, L) (record 0 never written,
// taken as zeros; record lmd = get_wds's for lmd) sorted by density,
// descending, stably for L <= 16; the first holds the answer, so lmd = L
// takes no part (CoreIndex.cpp:474).
int densest_coreindex_impl(hg_graph* g, double beta, hg_wds* answer, hg_wds* per_lmd) {
    HIP_TRY(hipSetDevice(g->device));
    const uint32_t L = g->L;
    const float fbeta = float(beta);
    std::vector<double> pw(L);
    for (uint32_t l = 0; l < L; ++l) pw[l] = std::pow(double(l + 1), double(fbeta));
    if (!g->hist) {
        hipError_t e = dalloc(&g->hist, size_t(L) * g->kcap);
        if (e != hipSuccess) return fail(HG_ERR_NOMEM, "densest histogram allocation failed");
    }
    const uint32_t grid = uint32_t(std::min<uint64_t>((g->nchunks * kChunk / 64 + kHistBlock - 1) / kHistBlock,
                                                       uint64_t(g->cus) * 8));
    std::vector<hg_wds> rec(size_t(L) + 1);
    for (auto& r : rec) std::memset(&r, 0, sizeof r);
    std::vector<uint64_t> h;
    hgc::Staging io(g->stream);
    for (uint32_t i = 0; i < L; ++i) {
        const int32_t km = g->h_kmax[i];
        hg_wds& r = rec[size_t(i) + 1];
        if (km < 1) continue;
        const uint32_t kc = uint32_t(km) + 1;
        HIP_TRY(hipMemsetAsync(g->hist, 0, sizeof(uint64_t) * size_t(L) * kc, g->stream));
        hipLaunchKernelGGL(k_edge_hist, dim3(std::max(grid, 1u)), dim3(kHistBlock), 0, g->stream, g->insts[i], g->chunks,
                           g->nchunks, g->hist, kc);
        HIP_TRY(hipGetLastError());
        h.assign(size_t(L) * kc, 0);
        HIP_TRY(io.d2h(h.data(), g->hist, sizeof(uint64_t) * h.size()));
        const std::vector<int64_t> len = cell_lengths(g, i);
        std::vector<uint64_t> acc(L, 0);
        std::vector<float> dens(size_t(km) + 1, 0.0f);
        std::vector<int32_t> sel(size_t(km) + 1, 0);
        for (int32_t k = km; k >= 1; --k) {  // suffix sums over m >= k
            float deg[64];
            for (uint32_t l = 0; l < L; ++l) {
                acc[l] += h[size_t(l) * kc + size_t(k)];
                const float n_edge = acc[l] >= (1ull << 24) ? 16777216.0f : float(acc[l]);
                deg[l] = n_edge / 2;
            }
            std::sort(deg, deg + L, [](float a, float c) { return a > c; });
            const uint32_t n_node = uint32_t(len[size_t(k)]);
            float best = 0.0f;
            for (uint32_t l = 0; l < L; ++l) {
                const float mad = deg[l] / float(n_node);
                const float d = float(double(mad) * pw[l]);
                if (d >= best) {
                    best = d;
                    sel[size_t(k)] = int32_t(l + 1);
                }
            }
            dens[size_t(k)] = best;
        }
        for (int32_t k = 1; k <= km; ++k) {  // k ascending, >=, as get_wds's loop
            if (dens[size_t(k)] >= r.max_density) {
                r.max_density = dens[size_t(k)];
                r.res_k = uint32_t(k);
                r.res_lmd = i + 1;
                r.res_len = uint32_t(len[size_t(k)]);
                r.num_select_layer = sel[size_t(k)];
            }
        }
    }
    if (per_lmd)
        for (uint32_t i = 0; i < L; ++i) per_lmd[i] = rec[size_t(i) + 1];
    if (answer) {
        hg_wds best = rec[0];
        for (uint32_t i = 1; i < L; ++i)  // records [0, L)
            if (rec[i].max_density > best.max_density) best = rec[i];
        *answer = best;
    }
    return HG_OK;
}

}  // namespace

extern "C" {

const char* hg_last_error(void) { return hgc::g_err.c_str(); }

int hg_graph_create(const hg_layer_csr* layers, int32_t L, int64_t n, int32_t offset_bits, void* stream,
                    hg_graph** out) {
    if (!out) return fail(HG_ERR_ARG, "out is NULL");
    *out = nullptr;
    hg_graph* g = new (std::nothrow) hg_graph;
    if (!g) return fail(HG_ERR_NOMEM, "host allocation failed");
    int rc;
    try {
        rc = create_impl(layers, L, n, offset_bits, stream, g);
    } catch (const std::bad_alloc&) {
        rc = fail(HG_ERR_NOMEM, "host allocation failed in hg_graph_create");
    }
    if (rc != HG_OK) {
        release(g);
        return rc;
    }
    *out = g;
    return HG_OK;
}

int hg_decompose(hg_graph* g) {
    if (!g) return fail(HG_ERR_ARG, "graph is NULL");
    try {
        return decompose_impl(g);
    } catch (const std::bad_alloc&) {
        return fail(HG_ERR_NOMEM, "host allocation failed in hg_decompose");
    }
}

int hg_densest_noalloc(hg_graph* g, double beta, hg_wds* out) {
    if (!g) return fail(HG_ERR_ARG, "graph is NULL");
    if (!g->decomposed) return fail(HG_ERR_STATE, "hg_decompose has not completed");
    if (!std::isfinite(beta)) return fail(HG_ERR_ARG, "beta must be finite");
    try {
        return densest_noalloc_impl(g, beta, out);
    } catch (const std::bad_alloc&) {
        return fail(HG_ERR_NOMEM, "host allocation failed in hg_densest_noalloc");
    }
}

int hg_set_phase_timing(hg_graph* g, int32_t on) {
    if (!g) return fail(HG_ERR_ARG, "graph is NULL");
    g->phase_timing = on != 0;
    return HG_OK;
}

int hg_core_index(hg_graph* g, int32_t* core_index) {
    if (!g || !core_index) return fail(HG_ERR_ARG, "NULL argument");
    if (!g->decomposed) return fail(HG_ERR_STATE, "hg_decompose has not completed");
    HIP_TRY(hipSetDevice(g->device));
    hgc::Staging io(g->stream);
    HIP_TRY(io.d2h(core_index, g->out, sizeof(int32_t) * size_t(g->n) * g->L));
    return HG_OK;
}

int hg_core_index_device(hg_graph* g, const int32_t** dev_ptr) {
    if (!g || !dev_ptr) return fail(HG_ERR_ARG, "NULL argument");
    if (!g->decomposed) return fail(HG_ERR_STATE, "hg_decompose has not completed");
    *dev_ptr = g->out;
    return HG_OK;
}

int hg_cells(hg_graph* g, int64_t* cell_len, int32_t cap, int32_t* kmax) {
    if (!g) return fail(HG_ERR_ARG, "graph is NULL");
    if (cell_len && cap < 1) return fail(HG_ERR_ARG, "cap must be >= 1");
    if (!g->decomposed) return fail(HG_ERR_STATE, "hg_decompose has not completed");
    for (uint32_t i = 0; i < g->L; ++i) {
        if (kmax) kmax[i] = g->h_kmax[i];
        if (cell_len) {
            const std::vector<int64_t> len = cell_lengths(g, i);
            for (int32_t k = 0; k < cap; ++k)
                cell_len[size_t(i) * cap + size_t(k)] = k < int32_t(len.size()) ? len[size_t(k)] : 0;
        }
    }
    return HG_OK;
}

int hg_shells(hg_graph* g, int32_t kcap, int64_t* shell_off, uint32_t* shell_vtx) {
    if (!g || !shell_off || !shell_vtx) return fail(HG_ERR_ARG, "NULL argument");
    if (!g->decomposed) return fail(HG_ERR_STATE, "hg_decompose has not completed");
    for (uint32_t i = 0; i < g->L; ++i)
        if (g->h_kmax[i] >= kcap) return fail(HG_ERR_ARG, "kcap must exceed every kmax");
    try {
        std::vector<int32_t> ci(size_t(g->n) * g->L);
        int rc = hg_core_index(g, ci.data());
        if (rc) return rc;
        const size_t stride = size_t(kcap) + 1;
        std::vector<int64_t> cnt(stride * g->L + 1, 0);
        for (uint32_t i = 0; i < g->L; ++i)
            for (uint32_t v = 0; v < g->n; ++v) {
                const int32_t x = ci[size_t(i) * g->n + v];
                if (x < 0 || x >= kcap) return fail(HG_ERR_DEVICE, "core index outside [0, kcap) in the result");
                cnt[i * stride + size_t(x) + 1]++;
            }
        shell_off[0] = 0;
        for (size_t t = 0; t < stride * g->L; ++t) shell_off[t + 1] = shell_off[t] + cnt[t + 1];
        std::vector<int64_t> pos(shell_off, shell_off + stride * g->L);
        for (uint32_t i = 0; i < g->L; ++i)
            for (uint32_t v = 0; v < g->n; ++v) shell_vtx[pos[i * stride + size_t(ci[size_t(i) * g->n + v])]++] = v;
    } catch (const std::bad_alloc&) {
        return fail(HG_ERR_NOMEM, "host allocation failed in hg_shells");
    }
    return HG_OK;
}

int hg_densest(hg_graph* g, double beta, hg_wds* out) {
    if (!g) return fail(HG_ERR_ARG, "graph is NULL");
    if (!g->decomposed) return fail(HG_ERR_STATE, "hg_decompose has not completed");
    if (!std::isfinite(beta)) return fail(HG_ERR_ARG, "beta must be finite");
    try {
        return densest_impl(g, beta, out);
    } catch (const std::bad_alloc&) {
        return fail(HG_ERR_NOMEM, "host allocation failed in hg_densest");
    }
}

int hg_densest_coreindex(hg_graph* g, double beta, hg_wds* answer, hg_wds* per_lmd) {
    if (!g) return fail(HG_ERR_ARG, "graph is NULL");
    if (!g->decomposed) return fail(HG_ERR_STATE, "hg_decompose has not completed");
    if (!std::isfinite(beta)) return fail(HG_ERR_ARG, "beta must be finite");
    try {
        return densest_coreindex_impl(g, beta, answer, per_lmd);
    } catch (const std::bad_alloc&) {
        return fail(HG_ERR_NOMEM, "host allocation failed in hg_densest_coreindex");
    }
}

int hg_cell_density(hg_graph* g, double* density, int32_t* num_select_layer, int32_t cap) {
    if (!g) return fail(HG_ERR_ARG, "graph is NULL");
    if (!g->densest_done) return fail(HG_ERR_STATE, "hg_densest has not completed");
    if (cap < 1) return fail(HG_ERR_ARG, "cap must be >= 1");
    for (uint32_t i = 0; i < g->L; ++i)
        for (int32_t k = 0; k < cap; ++k) {
            const bool in = k <= g->h_kmax[i] && uint32_t(k) < g->kcap;
            if (density) density[size_t(i) * cap + size_t(k)] = in ? g->h_density[size_t(i) * g->kcap + size_t(k)] : 0.0;
            if (num_select_layer)
                num_select_layer[size_t(i) * cap + size_t(k)] = in ? g->h_sel[size_t(i) * g->kcap + size_t(k)] : 0;
        }
    return HG_OK;
}

int hg_densest_vertices(hg_graph* g, uint32_t* out, int64_t cap, int64_t* count) {
    if (!g || !count) return fail(HG_ERR_ARG, "NULL argument");
    if (!g->densest_done) return fail(HG_ERR_STATE, "hg_densest has not completed");
    *count = 0;
    if (g->wds.res_lmd == 0) return HG_OK;
    try {
        std::vector<int32_t> row(g->n);
        HIP_TRY(hipSetDevice(g->device));
        hgc::Staging io(g->stream);
        HIP_TRY(io.d2h(row.data(), g->out + size_t(g->wds.res_lmd - 1) * g->n, sizeof(int32_t) * g->n));
        int64_t c = 0;
        for (uint32_t v = 0; v < g->n; ++v) {
            if (row[v] >= int32_t(g->wds.res_k)) {
                if (out && c < cap) out[c] = v;
                ++c;
            }
        }
        *count = c;
        if (out && c > cap) return fail(HG_ERR_ARG, "output capacity too small");
    } catch (const std::bad_alloc&) {
        return fail(HG_ERR_NOMEM, "host allocation failed in hg_densest_vertices");
    }
    return HG_OK;
}

int hg_last_timing(hg_graph* g, hg_timing* out) {
    if (!g || !out) return fail(HG_ERR_ARG, "NULL argument");
    *out = g->timing;
    return HG_OK;
}

int hg_graph_info(hg_graph* g, int64_t* n, int32_t* L, int64_t* directed_entries, int64_t* device_bytes) {
    if (!g) return fail(HG_ERR_ARG, "graph is NULL");
    if (n) *n = g->n;
    if (L) *L = int32_t(g->L);
    if (directed_entries) *directed_entries = int64_t(g->D);
    if (device_bytes) *device_bytes = g->device_bytes + g->cmp_bytes;
    return HG_OK;
}

int hg_graph_destroy(hg_graph* g) {
    if (!g) return fail(HG_ERR_ARG, "graph is NULL");
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    else (void)hipDeviceSynchronize();
    release(g);
    return HG_OK;
}

}  // extern "C"
