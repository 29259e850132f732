// hipgrid_main.cpp -- command-line driver with the reference's flags
// (main.cpp:50-105): -d dataset, -m method, -o layer order, -k/-lmd query,
// -num_thread, -beta; plus -root <dir> (default "../dataset", the
// reference's relative dataset root, main.cpp:98) and -phase 1 (times the
// scans and cascades of the decomposition separately and prints them with
// the device's busy share, the analogue of the reference's U_overall line,
// FCSyncLeft.cpp:240-245).
//
// Methods (all run on the GPU through include/hipgrid.h):
//   HipGrid     fcgrid decomposition; replaces -m ParGrid / ParGridPlus /
//               ParGridright (main.cpp:229-328).  With -k K -lmd L also writes
//               the cell's vertex list like saveCoreToLocal (Core/FCTree.cpp:
//               138-153, main.cpp:213-223).
//   HipGridMem  decomposition plus k-shells per cell; replaces -m CoreMem
//               (main.cpp:331-345).
//   HipGridWds  weighted densest subgraph; replaces -m WdsList
//   HipGridWdsCoreIndex  the same cells scored as -m WdsCoreIndex does
//               (CoreIndexParallel/CoreIndex.cpp:307-478, main.cpp:444-451)
//               (main.cpp:402-442).
// Output lines follow the reference's (method/dataset line, layer order,
// statistics, elapsed time, "mem = .. MB").
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hg_host.h"
#include "hipgrid.h"

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int die(const char* what, const char* msg) {
    std::fprintf(stderr, "%s failed: %s\n", what, msg);
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    std::string dataset = "example", method = "HipGrid", root = "../dataset";
    int order = 0;
    unsigned k = 0, lmd = 0, num_thread = 1;
    int phase = 0;
    double beta = 1.0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        const bool has = i + 1 < argc;
        if (a == "-d" && has) dataset = argv[++i];
        else if (a == "-m" && has) method = argv[++i];
        else if (a == "-o" && has) order = std::atoi(argv[++i]);
        else if (a == "-k" && has) k = unsigned(std::atoi(argv[++i]));
        else if (a == "-lmd" && has) lmd = unsigned(std::atoi(argv[++i]));
        else if (a == "-num_thread" && has) num_thread = unsigned(std::atoi(argv[++i]));
        else if (a == "-beta" && has) beta = std::atof(argv[++i]);
        else if (a == "-root" && has) root = argv[++i];
        else if (a == "-phase" && has) phase = std::atoi(argv[++i]);
    }
    if (method != "HipGrid" && method != "HipGridMem" && method != "HipGridWds" && method != "HipGridWdsCoreIndex") {
        std::fprintf(stderr, "unknown method %s (HipGrid, HipGridMem, HipGridWds, HipGridWdsCoreIndex)\n",
                     method.c_str());
        return 2;
    }
    std::printf("method = %s dataset = %s\n", method.c_str(), dataset.c_str());
    const std::string dir = root + "/" + dataset + "/";
    hgh_graph* file_graph = nullptr;
    if (hgh_load_text(dir.c_str(), 1, &file_graph) != HGH_OK) return die("load", hgh_last_error());
    const int L = hgh_graph_layers(file_graph);
    const int64_t n = hgh_graph_n(file_graph);
    std::vector<int32_t> ord(static_cast<size_t>(L), 0);
    hgh_layer_order(file_graph, order, ord.data());
    hgh_graph* g = nullptr;
    if (hgh_permute_layers(file_graph, ord.data(), &g) != HGH_OK) return die("order", hgh_last_error());
    for (int i = 0; i < L; ++i) std::printf("%d ", ord[size_t(i)]);
    std::printf("\n");
    char stats[256];
    hgh_statistics(g, stats, sizeof stats);
    std::printf("%s\n", stats);
    std::printf("#thread = %u (host threads are not used by the GPU path)\n", num_thread);

    std::vector<hg_layer_csr> layers(static_cast<size_t>(L));
    for (int i = 0; i < L; ++i) {
        layers[size_t(i)].row_ptr = hgh_graph_row_ptr(g, i);
        layers[size_t(i)].col = hgh_graph_col(g, i);
    }
    const double t0 = now_s();
    hg_graph* eng = nullptr;
    if (hg_graph_create(layers.data(), L, n, 64, nullptr, &eng) != HG_OK) return die("hg_graph_create", hg_last_error());
    const double t1 = now_s();
    if (phase) hg_set_phase_timing(eng, 1);
    if (hg_decompose(eng) != HG_OK) return die("hg_decompose", hg_last_error());
    const double t2 = now_s();
    hg_timing tm{};
    hg_last_timing(eng, &tm);
    std::vector<int32_t> kmax(static_cast<size_t>(L));
    hg_cells(eng, nullptr, 0, kmax.data());
    int64_t cells = 0;
    for (int32_t x : kmax) cells += x;
    int64_t D = 0;
    hg_graph_info(eng, nullptr, nullptr, &D, nullptr);
    std::printf("upload + layout time: %g seconds\n", t1 - t0);
    std::printf("HipGrid Elapsed time: %g seconds\n", t2 - t1);
    std::printf("device decomposition time: %g ms, levels = %lld, GTEPS = %g\n", tm.decompose_ms,
                (long long)tm.levels, double(L) * double(D) / (tm.decompose_ms * 1e-3) / 1e9);
    std::printf("count = %lld\n", (long long)cells);
    if (phase) {
        // kernels of the level loop over the device time of the decomposition
        // (the reference's U_overall is the peel threads' busy share)
        const double busy = tm.init_ms + tm.scan_ms + tm.cascade_ms;
        std::printf("level 0: %g ms, scans: %g ms, cascades: %g ms\n", tm.init_ms, tm.scan_ms, tm.cascade_ms);
        std::printf("U_overall = %g %% (device busy in level kernels)\n", 100.0 * busy / tm.decompose_ms);
    }

    if (method == "HipGridMem") {
        int32_t kc = 1;
        for (int32_t x : kmax) kc = std::max(kc, x + 1);
        std::vector<int64_t> off(size_t(L) * size_t(kc + 1) + 1);
        std::vector<uint32_t> vtx(size_t(n) * size_t(L));
        if (hg_shells(eng, kc, off.data(), vtx.data()) != HG_OK) return die("hg_shells", hg_last_error());
        for (int i = 0; i < L; ++i)
            for (int32_t kk = 1; kk <= kmax[size_t(i)]; ++kk) {
                const size_t s = size_t(i) * size_t(kc + 1) + size_t(kk);
                std::printf("shell k = %d lmd = %d size = %lld\n", kk, i + 1, (long long)(off[s + 1] - off[s]));
            }
    }
    if (method == "HipGridWds") {
        hg_wds w{};
        if (hg_densest(eng, beta, &w) != HG_OK) return die("hg_densest", hg_last_error());
        hg_last_timing(eng, &tm);
        std::printf("maximum_density = %g\n", double(w.max_density));
        std::printf("res_k = %u\n", w.res_k);
        std::printf("res_lmd = %u\n", w.res_lmd);
        std::printf("res_len = %u\n", w.res_len);
        std::printf("Wds device time: %g ms\n", tm.densest_ms);
    }
    if (method == "HipGridWdsCoreIndex") {
        hg_wds w{};
        const auto t0 = std::chrono::steady_clock::now();
        if (hg_densest_coreindex(eng, beta, &w, nullptr) != HG_OK) return die("hg_densest_coreindex", hg_last_error());
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("maximum_density = %g\n", double(w.max_density));
        std::printf("res_k = %u\n", w.res_k);
        std::printf("res_lmd = %u\n", w.res_lmd);
        std::printf("Wds CoreIndex Time = %g\n", secs);
    }
    if (k != 0 && lmd != 0 && method == "HipGrid") {
        std::vector<int32_t> ci(size_t(n) * size_t(L));
        hg_core_index(eng, ci.data());
        std::printf("%u %u\n", k, lmd);
        if (lmd > unsigned(L) || int32_t(k) > kmax[lmd - 1]) {
            std::printf("No statisfy result\n");
        } else {
            int64_t len = 0;
            for (int64_t v = 0; v < n; ++v) len += ci[size_t(lmd - 1) * size_t(n) + size_t(v)] >= int32_t(k);
            const std::string res = "../output/" + dataset + "_k_" + std::to_string(k) + "_lmd_" + std::to_string(lmd) +
                                    "_HipGrid.txt";
            std::printf("res_file = %s\n", res.c_str());
            FILE* f = std::fopen(res.c_str(), "w");
            if (f) {
                std::fprintf(f, "########## k_%u_lmd_%u_length_%lld\n", k, lmd, (long long)len);
                for (int64_t v = 0; v < n; ++v)
                    if (ci[size_t(lmd - 1) * size_t(n) + size_t(v)] >= int32_t(k))
                        std::fprintf(f, "%llu\n", (unsigned long long)hgh_graph_label(g, v));
                std::fclose(f);
            }
        }
    }
    std::printf("mem = %g MB\n", hgh_peak_rss_mb());
    hg_graph_destroy(eng);
    hgh_graph_free(g);
    hgh_graph_free(file_graph);
    return 0;
}
