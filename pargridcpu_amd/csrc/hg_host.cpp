// hg_host.cpp -- host-side graph formats for the HipGrid FirmCore engine:
// the seeded Chung-Lu multilayer generator, the symmetric per-layer CSR
// builder, and the reference's text format (mlg.conf + edge lists) reader and
// writer.  Plain C ABI declared in include/hg_host.h.
//
// The text reader follows the reference loader's observable behaviour
// (Graph/MultilayerGraph.cpp:47-166, Graph/Graph.cpp:17-50): ids in
// first-appearance order over the layers in mlg.conf order, self-loops
// dropped, duplicates merged, rows sorted ascending.  It is our own code: a
// buffered tokenizer plus a counting-sort CSR build instead of an
// unordered_map-fed edge array and std::sort over (s, t) pairs.

#include "hg_host.h"

#include <omp.h>
#include <sys/resource.h>
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// Uninitialised heap array (avoids zero-fill of multi-GB buffers).
template <typename T>
struct Buf {
    T* p = nullptr;
    size_t n = 0;
    Buf() = default;
    explicit Buf(size_t count) { reset(count); }
    ~Buf() { std::free(p); }
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    Buf(Buf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    Buf& operator=(Buf&& o) noexcept {
        if (this != &o) { std::free(p); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    void reset(size_t count) {
        std::free(p);
        p = nullptr;
        n = count;
        if (count) {
            p = static_cast<T*>(std::malloc(count * sizeof(T)));
            if (!p) throw std::bad_alloc();
        }
    }
    T& operator[](size_t i) { return p[i]; }
    const T& operator[](size_t i) const { return p[i]; }
};

// ---------------------------------------------------------------- SplitMix64
// Output number t (t = 0, 1, ...) of a SplitMix64 stream seeded with s is
// mix(s + (t + 1) * golden); computing it directly makes every draw
// addressable, so parallel loops give thread-count independent results.
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t draw(uint64_t stream, uint64_t t) { return mix64(stream + (t + 1) * kGolden); }

enum Purpose : uint64_t { kSelect = 1, kPermute = 2, kEdges = 3, kEdgeCount = 4 };

// Seed of the substream for (layer, purpose).
inline uint64_t substream(uint64_t seed, int layer, uint64_t purpose) {
    return mix64(seed ^ mix64((uint64_t(layer) << 8) + purpose + 0x5851F42D4C957F2Dull));
}

// Uniform integer in [0, bound) from a 64-bit draw (multiply-high).
inline uint64_t below(uint64_t r, uint64_t bound) {
    return uint64_t((unsigned __int128)r * bound >> 64);
}

inline double unit_double(uint64_t r) { return double(r >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

struct hgh_graph {
    int64_t n = 0;
    int32_t L = 0;
    std::vector<Buf<int64_t>> row_ptr;  // L x (n+1)
    std::vector<Buf<uint32_t>> col;     // L x m_l
    std::vector<uint64_t> labels;       // empty: label(v) = v
};

namespace {

// Builds a symmetric, sorted, de-duplicated CSR row set from undirected pairs
// (a[j], b[j]), j < m; pairs with a == b are dropped.
void build_symmetric_csr(int64_t n, const uint32_t* a, const uint32_t* b, int64_t m,
                         Buf<int64_t>& row_ptr, Buf<uint32_t>& col) {
    std::unique_ptr<std::atomic<uint32_t>[]> cnt(new std::atomic<uint32_t>[n]);
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < n; ++v) cnt[v].store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < m; ++j) {
        if (a[j] == b[j]) continue;
        cnt[a[j]].fetch_add(1, std::memory_order_relaxed);
        cnt[b[j]].fetch_add(1, std::memory_order_relaxed);
    }
    Buf<int64_t> off(size_t(n) + 1);
    off[0] = 0;
    for (int64_t v = 0; v < n; ++v) off[v + 1] = off[v] + cnt[v].load(std::memory_order_relaxed);
    const int64_t total = off[n];
    Buf<uint32_t> tmp{size_t(total)};
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < n; ++v) cnt[v].store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < m; ++j) {
        const uint32_t x = a[j], y = b[j];
        if (x == y) continue;
        tmp[off[x] + cnt[x].fetch_add(1, std::memory_order_relaxed)] = y;
        tmp[off[y] + cnt[y].fetch_add(1, std::memory_order_relaxed)] = x;
    }
    // Sort + unique each row in place; record the unique length.
    Buf<uint32_t> ulen{size_t(n)};
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t v = 0; v < n; ++v) {
        uint32_t* s = tmp.p + off[v];
        uint32_t* e = tmp.p + off[v + 1];
        std::sort(s, e);
        ulen[v] = uint32_t(std::unique(s, e) - s);
    }
    row_ptr.reset(size_t(n) + 1);
    row_ptr[0] = 0;
    for (int64_t v = 0; v < n; ++v) row_ptr[v + 1] = row_ptr[v] + ulen[v];
    col.reset(size_t(row_ptr[n]));
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t v = 0; v < n; ++v)
        std::memcpy(col.p + row_ptr[v], tmp.p + off[v], sizeof(uint32_t) * ulen[v]);
}

// Vose alias table over weights w[0..n), quantised to 32-bit thresholds so
// that sampling is pure integer arithmetic.  A draw r picks column
// c = below(r >> 32 ...), then keeps c if (uint32)r < thr[c], else alias[c].
void build_alias(const double* w, int64_t n, Buf<uint32_t>& thr, Buf<uint32_t>& alias) {
    thr.reset(size_t(n));
    alias.reset(size_t(n));
    double sum = 0.0;
    for (int64_t i = 0; i < n; ++i) sum += w[i];
    Buf<double> p{size_t(n)};
    Buf<uint32_t> small{size_t(n)}, large{size_t(n)};
    int64_t ns = 0, nl = 0;
    const double scale = double(n) / sum;
    for (int64_t i = 0; i < n; ++i) {
        p[i] = w[i] * scale;
        if (p[i] < 1.0) small[ns++] = uint32_t(i);
        else large[nl++] = uint32_t(i);
    }
    auto quant = [](double x) -> uint32_t {
        if (x >= 1.0) return 0xFFFFFFFFu;
        if (x <= 0.0) return 0u;
        double q = std::floor(x * 4294967296.0);
        return q >= 4294967295.0 ? 0xFFFFFFFFu : uint32_t(q);
    };
    while (ns > 0 && nl > 0) {
        const uint32_t s = small[--ns];
        const uint32_t g = large[nl - 1];
        thr[s] = quant(p[s]);
        alias[s] = g;
        p[g] = (p[g] + p[s]) - 1.0;
        if (p[g] < 1.0) {
            --nl;
            small[ns++] = g;
        }
    }
    while (nl > 0) {
        const uint32_t g = large[--nl];
        thr[g] = 0xFFFFFFFFu;
        alias[g] = g;
    }
    while (ns > 0) {  // round-off leftovers
        const uint32_t s = small[--ns];
        thr[s] = 0xFFFFFFFFu;
        alias[s] = s;
    }
}

inline uint32_t alias_pick(uint64_t r, int64_t n, const uint32_t* thr, const uint32_t* alias) {
    const uint32_t c = uint32_t((uint64_t(uint32_t(r >> 32)) * uint64_t(n)) >> 32);
    return uint32_t(r) < thr[c] ? c : alias[c];
}

int64_t layer_edge_count(const hgh_gen_params& p, int layer) {
    if (p.m_lo == p.m_hi) return p.m_lo;
    const double u = unit_double(draw(substream(p.seed, layer, kEdgeCount), 0));
    const double lo = std::log(double(p.m_lo)), hi = std::log(double(p.m_hi));
    return int64_t(std::floor(std::exp(lo + u * (hi - lo))));
}

// Per-layer weights: base weights with floor(rho*n) positions permuted among
// themselves, then the optional expected-degree cap.
void layer_weights(const hgh_gen_params& p, int layer, const double* base, int64_t m, Buf<double>& w) {
    const int64_t n = p.n;
    w.reset(size_t(n));
    std::memcpy(w.p, base, sizeof(double) * size_t(n));
    const int64_t s = int64_t(std::floor(p.rho * double(n)));
    if (s >= 2) {
        // Partial Fisher-Yates: idx[0..s) becomes a uniform random s-subset.
        Buf<uint32_t> idx{size_t(n)};
        for (int64_t i = 0; i < n; ++i) idx[i] = uint32_t(i);
        const uint64_t sel = substream(p.seed, layer, kSelect);
        for (int64_t i = 0; i < s && i < n - 1; ++i) {
            const int64_t j = i + int64_t(below(draw(sel, uint64_t(i)), uint64_t(n - i)));
            std::swap(idx[i], idx[j]);
        }
        // Full Fisher-Yates over the selected positions' weights.
        Buf<double> vals{size_t(s)};
        for (int64_t i = 0; i < s; ++i) vals[i] = w[idx[i]];
        const uint64_t per = substream(p.seed, layer, kPermute);
        for (int64_t i = s - 1; i >= 1; --i) {
            const int64_t j = int64_t(below(draw(per, uint64_t(i)), uint64_t(i + 1)));
            std::swap(vals[i], vals[j]);
        }
        for (int64_t i = 0; i < s; ++i) w[idx[i]] = vals[i];
    }
    if (p.cap_deg > 0.0 && m > 0) {
        double sum = 0.0;
        for (int64_t i = 0; i < n; ++i) sum += w[i];
        const double cap_w = p.cap_deg * sum / (2.0 * double(m));
        for (int64_t i = 0; i < n; ++i) w[i] = std::min(w[i], cap_w);
    }
}

}  // namespace

extern "C" {

const char* hgh_last_error(void) { return g_err.c_str(); }

int hgh_generate(const hgh_gen_params* p, hgh_graph** out) {
    if (!p || !out) return fail(HGH_ERR_ARG, "null argument");
    *out = nullptr;
    if (p->n <= 1 || p->n > int64_t(0xFFFFFFFFll) || p->L <= 0 || p->L > 1024)
        return fail(HGH_ERR_RANGE, "n must be in [2, 2^32) and L in [1, 1024]");
    if (!(p->gamma > 1.0) || p->rho < 0.0 || p->rho > 1.0 || p->m_lo < 0 || p->m_hi < p->m_lo)
        return fail(HGH_ERR_ARG, "bad gamma/rho/m range");
    try {
        std::unique_ptr<hgh_graph> g(new hgh_graph);
        const int64_t n = p->n;
        const int L = p->L;
        g->n = n;
        g->L = L;
        g->row_ptr.resize(size_t(L));
        g->col.resize(size_t(L));
        Buf<double> base{size_t(n)};
        const double expo = -1.0 / (p->gamma - 1.0);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) base[i] = std::pow(double(i + 1), expo);

        std::vector<int64_t> m(static_cast<size_t>(L));
        for (int l = 0; l < L; ++l) m[size_t(l)] = layer_edge_count(*p, l);

        // Alias tables: the per-layer set-up is sequential work, so layers
        // run side by side (bounded to keep peak memory moderate).
        std::vector<Buf<uint32_t>> thr(static_cast<size_t>(L)), alias(static_cast<size_t>(L));
        const int par = std::max(1, std::min(L, omp_get_max_threads()));
#pragma omp parallel for schedule(dynamic, 1) num_threads(par)
        for (int l = 0; l < L; ++l) {
            Buf<double> w;
            layer_weights(*p, l, base.p, m[size_t(l)], w);
            build_alias(w.p, n, thr[size_t(l)], alias[size_t(l)]);
        }
        base.reset(0);

        for (int l = 0; l < L; ++l) {
            const int64_t ml = m[size_t(l)];
            Buf<uint32_t> a{size_t(ml)}, b{size_t(ml)};
            const uint64_t es = substream(p->seed, l, kEdges);
            const uint32_t* T = thr[size_t(l)].p;
            const uint32_t* A = alias[size_t(l)].p;
#pragma omp parallel for schedule(static)
            for (int64_t j = 0; j < ml; ++j) {
                a[j] = alias_pick(draw(es, 2 * uint64_t(j)), n, T, A);
                b[j] = alias_pick(draw(es, 2 * uint64_t(j) + 1), n, T, A);
            }
            thr[size_t(l)].reset(0);
            alias[size_t(l)].reset(0);
            build_symmetric_csr(n, a.p, b.p, ml, g->row_ptr[size_t(l)], g->col[size_t(l)]);
        }
        *out = g.release();
        return HGH_OK;
    } catch (const std::bad_alloc&) {
        return fail(HGH_ERR_NOMEM, "out of host memory in hgh_generate");
    }
}

// ------------------------------------------------------------- text format

namespace {

// Buffered reader of unsigned decimal tokens.
struct TokenReader {
    FILE* f = nullptr;
    std::vector<char> buf;
    size_t pos = 0, len = 0;
    bool eof = false;
    explicit TokenReader(FILE* fp) : f(fp), buf(1 << 20) {}
    int getc_() {
        if (pos == len) {
            if (eof) return -1;
            len = std::fread(buf.data(), 1, buf.size(), f);
            pos = 0;
            if (len == 0) { eof = true; return -1; }
        }
        return (unsigned char)buf[pos++];
    }
    void ungetc_() { --pos; }
    // 1: parsed; 0: end of input or a token that is not an unsigned integer
    // (stream extraction stops there).
    int next(uint32_t& out) {
        int c;
        do { c = getc_(); } while (c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f');
        if (c < 0) return 0;
        if (c == '+') c = getc_();
        if (c < '0' || c > '9') return 0;
        uint64_t v = 0;
        while (c >= '0' && c <= '9') {
            v = v * 10 + uint64_t(c - '0');
            if (v > 0xFFFFFFFFull) v = 0xFFFFFFFFull;  // saturate like a failed extraction
            c = getc_();
        }
        if (c >= 0) ungetc_();
        out = uint32_t(v);
        return 1;
    }
};

bool read_conf(const std::string& dir, std::vector<std::string>& files) {
    FILE* f = std::fopen((dir + "mlg.conf").c_str(), "rb");
    if (!f) return false;
    std::string line;
    int c;
    while (true) {
        c = std::fgetc(f);
        if (c == EOF || c == '\n') {
            if (!line.empty()) files.push_back(line);
            line.clear();
            if (c == EOF) break;
        } else {
            line.push_back(char(c));
        }
    }
    std::fclose(f);
    return true;
}

}  // namespace

int hgh_load_text(const char* dir_c, int32_t write_vtx_map, hgh_graph** out) {
    if (!dir_c || !out) return fail(HGH_ERR_ARG, "null argument");
    *out = nullptr;
    std::string dir(dir_c);
    if (!dir.empty() && dir.back() != '/') dir.push_back('/');
    std::vector<std::string> files;
    if (!read_conf(dir, files)) return fail(HGH_ERR_IO, "cannot open " + dir + "mlg.conf");
    if (files.empty()) return fail(HGH_ERR_IO, "mlg.conf lists no layer");
    try {
        std::unordered_map<uint64_t, uint32_t> vtx2id;
        std::vector<uint64_t> labels;
        std::vector<std::vector<uint32_t>> ea(files.size()), eb(files.size());
        FILE* map_out = nullptr;
        if (write_vtx_map) {
            map_out = std::fopen((dir + "__vtx_map.txt").c_str(), "wb");
            if (!map_out) return fail(HGH_ERR_IO, "cannot write " + dir + "__vtx_map.txt");
        }
        auto id_of = [&](uint32_t label) -> uint32_t {
            auto it = vtx2id.find(label);
            if (it != vtx2id.end()) return it->second;
            const uint32_t id = uint32_t(labels.size());
            vtx2id.emplace(label, id);
            labels.push_back(label);
            if (map_out) std::fprintf(map_out, "%u %u\n", label, id);
            return id;
        };
        for (size_t l = 0; l < files.size(); ++l) {
            FILE* f = std::fopen((dir + files[l]).c_str(), "rb");
            if (!f) {
                if (map_out) std::fclose(map_out);
                return fail(HGH_ERR_IO, "cannot open layer file " + dir + files[l]);
            }
            TokenReader rd(f);
            uint32_t u = 0, v = 0;
            bool have_prev = false;
            while (true) {
                if (!rd.next(u)) break;
                // A lone trailing token pairs with the previous second
                // endpoint, as a failed ">> v" leaves v unchanged.
                if (!rd.next(v)) {
                    if (!have_prev) break;
                    if (u != v) {
                        const uint32_t iu = id_of(u), iv = id_of(v);
                        ea[l].push_back(iu);
                        eb[l].push_back(iv);
                    }
                    break;
                }
                have_prev = true;
                if (u == v) continue;
                const uint32_t iu = id_of(u), iv = id_of(v);
                ea[l].push_back(iu);
                eb[l].push_back(iv);
            }
            std::fclose(f);
        }
        if (map_out) std::fclose(map_out);
        std::unique_ptr<hgh_graph> g(new hgh_graph);
        g->n = int64_t(labels.size());
        g->L = int32_t(files.size());
        g->row_ptr.resize(files.size());
        g->col.resize(files.size());
        for (size_t l = 0; l < files.size(); ++l) {
            if (g->n == 0) {
                g->row_ptr[l].reset(1);
                g->row_ptr[l][0] = 0;
                continue;
            }
            build_symmetric_csr(g->n, ea[l].data(), eb[l].data(), int64_t(ea[l].size()), g->row_ptr[l], g->col[l]);
            std::vector<uint32_t>().swap(ea[l]);
            std::vector<uint32_t>().swap(eb[l]);
        }
        g->labels.swap(labels);
        *out = g.release();
        return HGH_OK;
    } catch (const std::bad_alloc&) {
        return fail(HGH_ERR_NOMEM, "out of host memory in hgh_load_text");
    }
}

int hgh_write_text(const hgh_graph* g, const char* dir_c) {
    if (!g || !dir_c) return fail(HGH_ERR_ARG, "null argument");
    std::string dir(dir_c);
    if (!dir.empty() && dir.back() != '/') dir.push_back('/');
    FILE* conf = std::fopen((dir + "mlg.conf").c_str(), "wb");
    if (!conf) return fail(HGH_ERR_IO, "cannot write " + dir + "mlg.conf");
    for (int l = 0; l < g->L; ++l) std::fprintf(conf, "layer_%d.txt\n", l);
    std::fclose(conf);
    std::vector<char> out;
    out.reserve(1 << 22);
    char tmp[48];
    for (int l = 0; l < g->L; ++l) {
        FILE* f = std::fopen((dir + "layer_" + std::to_string(l) + ".txt").c_str(), "wb");
        if (!f) return fail(HGH_ERR_IO, "cannot write layer file");
        const int64_t* rp = g->row_ptr[size_t(l)].p;
        const uint32_t* c = g->col[size_t(l)].p;
        for (int64_t v = 0; v < g->n; ++v) {
            for (int64_t i = rp[v]; i < rp[v + 1]; ++i) {
                const uint32_t u = c[i];
                if (u <= uint32_t(v)) continue;
                const int k = std::snprintf(tmp, sizeof tmp, "%llu %llu\n",
                                            (unsigned long long)hgh_graph_label(g, v),
                                            (unsigned long long)hgh_graph_label(g, u));
                out.insert(out.end(), tmp, tmp + k);
                if (out.size() > (1 << 22)) {
                    std::fwrite(out.data(), 1, out.size(), f);
                    out.clear();
                }
            }
        }
        std::fwrite(out.data(), 1, out.size(), f);
        out.clear();
        std::fclose(f);
    }
    return HGH_OK;
}

int hgh_layer_order(const hgh_graph* g, int32_t o, int32_t* order) {
    if (!g || !order) return fail(HGH_ERR_ARG, "null argument");
    const int L = g->L;
    std::vector<uint32_t> ord(static_cast<size_t>(L));
    for (int i = 0; i < L; ++i) ord[size_t(i)] = uint32_t(i);
    if (o == 1 || o == 2) {
        std::vector<float> density(static_cast<size_t>(L));
        for (int i = 0; i < L; ++i) {
            const uint32_t m = uint32_t(hgh_graph_entries(g, i));
            density[size_t(i)] = float(m >> 1) / float(uint32_t(g->n));
        }
        if (o == 1)
            std::sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return density[x] < density[y]; });
        else
            std::sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return density[x] > density[y]; });
    } else if (o == 3 || o == 4) {
        std::vector<uint32_t> ne(static_cast<size_t>(L));
        for (int i = 0; i < L; ++i) ne[size_t(i)] = uint32_t(hgh_graph_entries(g, i));
        if (o == 3)
            std::sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return ne[x] < ne[y]; });
        else
            std::sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return ne[x] > ne[y]; });
    }
    for (int i = 0; i < L; ++i) order[i] = int32_t(ord[size_t(i)]);
    return HGH_OK;
}

int hgh_permute_layers(const hgh_graph* g, const int32_t* order, hgh_graph** out) {
    if (!g || !order || !out) return fail(HGH_ERR_ARG, "null argument");
    *out = nullptr;
    std::vector<int> seen(size_t(g->L), 0);
    for (int i = 0; i < g->L; ++i) {
        if (order[i] < 0 || order[i] >= g->L || seen[size_t(order[i])]++)
            return fail(HGH_ERR_ARG, "order is not a permutation of 0..L-1");
    }
    try {
        std::unique_ptr<hgh_graph> h(new hgh_graph);
        h->n = g->n;
        h->L = g->L;
        h->labels = g->labels;
        h->row_ptr.resize(size_t(g->L));
        h->col.resize(size_t(g->L));
        for (int i = 0; i < g->L; ++i) {
            const auto& rp = g->row_ptr[size_t(order[i])];
            const auto& c = g->col[size_t(order[i])];
            h->row_ptr[size_t(i)].reset(rp.n);
            if (rp.n) std::memcpy(h->row_ptr[size_t(i)].p, rp.p, sizeof(int64_t) * rp.n);
            h->col[size_t(i)].reset(c.n);
            if (c.n) std::memcpy(h->col[size_t(i)].p, c.p, sizeof(uint32_t) * c.n);
        }
        *out = h.release();
        return HGH_OK;
    } catch (const std::bad_alloc&) {
        return fail(HGH_ERR_NOMEM, "out of host memory in hgh_permute_layers");
    }
}

int64_t hgh_graph_n(const hgh_graph* g) { return g ? g->n : -1; }
int32_t hgh_graph_layers(const hgh_graph* g) { return g ? g->L : -1; }
const int64_t* hgh_graph_row_ptr(const hgh_graph* g, int32_t l) {
    return (g && l >= 0 && l < g->L) ? g->row_ptr[size_t(l)].p : nullptr;
}
const uint32_t* hgh_graph_col(const hgh_graph* g, int32_t l) {
    return (g && l >= 0 && l < g->L) ? g->col[size_t(l)].p : nullptr;
}
int64_t hgh_graph_entries(const hgh_graph* g, int32_t l) {
    if (!g || l < 0 || l >= g->L) return -1;
    return g->row_ptr[size_t(l)].n ? g->row_ptr[size_t(l)][size_t(g->n)] : 0;
}
uint64_t hgh_graph_label(const hgh_graph* g, int64_t v) {
    if (!g || v < 0 || v >= g->n) return ~0ull;
    return g->labels.empty() ? uint64_t(v) : g->labels[size_t(v)];
}
void hgh_graph_free(hgh_graph* g) { delete g; }

int hgh_statistics(const hgh_graph* g, char* buf, int32_t cap) {
    if (!g || !buf || cap <= 0) return fail(HGH_ERR_ARG, "null argument");
    uint32_t max_m = 0, sum_m = 0;
    for (int l = 0; l < g->L; ++l) {
        const uint32_t m = uint32_t(hgh_graph_entries(g, l)) >> 1;
        max_m = std::max(max_m, m);
        sum_m += m;
    }
    std::snprintf(buf, size_t(cap), "|L| = %d, |V| = %lld, |E| = %u, max_|E_i| = %u", g->L,
                  (long long)g->n, sum_m, max_m);
    return HGH_OK;
}

double hgh_peak_rss_mb(void) {
    struct rusage ru;
    getrusage(RUSAGE_SELF, &ru);
    return double(ru.ru_maxrss) * 1024.0 / 1024.0 / 1024.0;
}

}  // extern "C"
