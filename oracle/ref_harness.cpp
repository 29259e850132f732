// ref_harness.cpp -- drives the reference's own FirmCore routines, compiled
// from the sources under the reference tree (oracle/Makefile), and dumps
// their results so tests/golden/make_golden.py can turn them into fixtures.
//
// TEST INFRASTRUCTURE ONLY: runs in the build container, never shipped in
// the product path.  The harness only calls reference entry points and walks
// the result trees they return; it computes nothing itself.
//
// usage: ref_harness <mode> <dataset_dir/> <out_prefix> [beta]
//   mode  lpe | kpe     FCTreeBuilderLeft/Right::Execute (Core/FCTreeBuilder*.cpp);
//                       every cell's core[] list -> per-vertex core index
//         coremem       FCSyncMem::Execute (CoreParallel/FCSyncMem.cpp:199-243);
//                       per-cell k-shells (diff_res / last valid_vertex) -> index
//         coreindex     get_index2 per lmd, as CoreIndex::Execute runs it
//                       (CoreIndexParallel/CoreIndex.cpp:62-109, 168-213)
//         pargrid | pargridplus | pargridright
//                       FCSyncLeft::Execute / ExecuteMix, FCSyncRight::Execute
//                       (CoreParallel/FCSync*.cpp); cell (k, lmd, length) list
//         wds           FCPathLevelLeft::Execute(mg, tree, true, beta) then
//                       FCTree::WeightDenestSubgraph (main.cpp:402-424)
//         wdscoreindex  get_wds per lmd (CoreIndexParallel/CoreIndex.cpp:347-413),
//                       then CoreIndex::WdsCoreIndex itself (:415-479), whose
//                       printed answer goes to stdout
// outputs  <out_prefix>.labels  uint64[n]  id2vtx (MultilayerGraph::LoadId2VtxMap)
//          <out_prefix>.index   int32[L][n] (index modes)
//          <out_prefix>.cells   text "lmd k length [density sel]" per cell
//          <out_prefix>.wds     text "max_density res_k res_lmd res_len" (wds)
//          <out_prefix>.wdsci   text "lmd density res_k res_lmd" per lmd (wdscoreindex)
#include "Graph/MultilayerGraph.h"
#include "Core/FCTree.h"
#include "Core/FCTreeBuilderLeft.h"
#include "Core/FCTreeBuilderRight.h"
#include "Core/FCPathLevelLeft.h"
#include "CoreParallel/FCCoreTree.h"
#include "CoreParallel/FCSyncLeft.h"
#include "CoreParallel/FCSyncRight.h"
#include "CoreParallel/FCSyncMem.h"
#include "CoreIndexParallel/CoreIndex.h"

#include <cstdio>
#include <cstdlib>
#include <vector>

// Defined with external linkage in CoreIndexParallel/CoreIndex.cpp:62.
uint* get_index2(MultilayerGraph& mg, uint** degs, uint lmd, ll_uint* id2vtx);

// The per-lmd result record of CoreIndex.cpp:9-13 (declared there, in the
// .cpp only), needed to call get_wds (CoreIndex.cpp:347, external linkage).
struct MaxDensity {
    uint resk;
    uint reslmd;
    float density;
};
uint* get_wds(MultilayerGraph& mg, uint** degs, uint lmd, ll_uint* id2vtx, MaxDensity& maximum_density, float beta,
              uint& res_k, uint& res_lmd);

namespace {

struct Cell {
    unsigned k, lmd, length;
    double density;
    int sel;
};

void walk_fctree(coreNode* node, std::vector<Cell>& cells, std::vector<int>& index, unsigned n,
                 bool with_lists) {
    if (node == nullptr || node->k == 0) return;
    cells.push_back({node->k, node->lmd, node->length, node->layer_density, node->num_select_layer});
    if (with_lists && node->core) {
        int* row = index.data() + size_t(node->lmd - 1) * n;
        for (unsigned i = 0; i < node->length; ++i) {
            const unsigned v = node->core[i];
            if (row[v] < int(node->k)) row[v] = int(node->k);
        }
    }
    walk_fctree(node->left, cells, index, n, with_lists);
    walk_fctree(node->right, cells, index, n, with_lists);
}

void walk_ptree(coreNodeP* node, std::vector<Cell>& cells) {
    if (node == nullptr || node->k == 0) return;
    cells.push_back({node->k, node->lmd, node->length, 0.0, 0});
    walk_ptree(node->left, cells);
    walk_ptree(node->right, cells);
}

void walk_memtree(coreNodeP* node, std::vector<Cell>& cells, std::vector<int>& index, unsigned n) {
    if (node == nullptr || node->k == 0) return;
    cells.push_back({node->k, node->lmd, node->length, 0.0, 0});
    const std::vector<uint>& shell = node->diff_res.empty() ? node->valid_vertex : node->diff_res;
    int* row = index.data() + size_t(node->lmd - 1) * n;
    for (uint v : shell) row[v] = int(node->k);
    walk_memtree(node->left, cells, index, n);
    walk_memtree(node->right, cells, index, n);
}

bool write_bin(const std::string& path, const void* p, size_t bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(p, 1, bytes, f) == bytes;
    std::fclose(f);
    return ok;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <mode> <dataset_dir/> <out_prefix> [beta]\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    const std::string dir = argv[2];
    const std::string out = argv[3];
    const double beta = argc > 4 ? std::atof(argv[4]) : 1.0;

    MultilayerGraph mg;
    mg.LoadFromFile(dir);
    const unsigned n = mg.GetN();
    const unsigned L = mg.getLayerNumber();
    std::vector<ll_uint> id2vtx(n);
    mg.LoadId2VtxMap(id2vtx.data());
    if (!write_bin(out + ".labels", id2vtx.data(), sizeof(ll_uint) * n)) return 3;

    std::vector<Cell> cells;
    std::vector<int> index(size_t(L) * n, 0);
    bool have_index = false;

    if (mode == "lpe" || mode == "kpe") {
        FCTree tree(1, 1, n);
        if (mode == "lpe") FCTreeBuilderLeft::Execute(mg, tree);
        else FCTreeBuilderRight::Execute(mg, tree);
        walk_fctree(tree.getNode(), cells, index, n, true);
        have_index = true;
    } else if (mode == "coremem") {
        FCCoreTree tree(1, 1, n);
        FCSyncMem::Execute(mg, tree);
        walk_memtree(tree.getNode(), cells, index, n);
        have_index = true;
    } else if (mode == "coreindex") {
        for (uint lmd = 1; lmd <= L; ++lmd) {
            uint** degs = new uint*[n];
            for (uint v = 0; v < n; ++v) {
                degs[v] = new uint[L];
                for (uint l = 0; l < L; ++l) degs[v][l] = mg.GetGraph(l).GetAdjLst()[v][0];
            }
            uint* iv = get_index2(mg, degs, lmd, id2vtx.data());
            for (uint v = 0; v < n; ++v) index[size_t(lmd - 1) * n + v] = int(iv[v]);
            delete[] iv;
            for (uint v = 0; v < n; ++v) delete[] degs[v];
            delete[] degs;
        }
        have_index = true;
    } else if (mode == "pargrid" || mode == "pargridplus" || mode == "pargridright") {
        FCCoreTree tree(1, 1, n);
        if (mode == "pargrid") FCSyncLeft::Execute(mg, tree);
        else if (mode == "pargridplus") FCSyncLeft::ExecuteMix(mg, tree);
        else FCSyncRight::Execute(mg, tree);
        walk_ptree(tree.getNode(), cells);
    } else if (mode == "wds") {
        FCTree tree(1, 1, n);
        FCPathLevelLeft::Execute(mg, tree, true, beta);
        float maximum_density = 0.0f;
        uint res_k = 0, res_lmd = 0, res_len = 0;
        tree.WeightDenestSubgraph(tree.getNode(), maximum_density, res_k, res_lmd, res_len);
        walk_fctree(tree.getNode(), cells, index, n, false);
        FILE* f = std::fopen((out + ".wds").c_str(), "w");
        if (!f) return 3;
        std::fprintf(f, "%.9g %u %u %u\n", double(maximum_density), res_k, res_lmd, res_len);
        std::fclose(f);
    } else if (mode == "wdscoreindex") {
        FILE* f = std::fopen((out + ".wdsci").c_str(), "w");
        if (!f) return 3;
        for (uint lmd = 1; lmd <= L; ++lmd) {
            uint** degs = new uint*[n];
            for (uint v = 0; v < n; ++v) {
                degs[v] = new uint[L];
                for (uint l = 0; l < L; ++l) degs[v][l] = mg.GetGraph(l).GetAdjLst()[v][0];
            }
            MaxDensity md{0, 0, 0.0f};
            uint rk = 0, rl = 0;
            uint* core = get_wds(mg, degs, lmd, id2vtx.data(), md, float(beta), rk, rl);
            std::fprintf(f, "%u %.9g %u %u\n", lmd, double(md.density), md.resk, md.reslmd);
            delete[] core;
            for (uint v = 0; v < n; ++v) delete[] degs[v];
            delete[] degs;
        }
        std::fclose(f);
        CoreIndex::WdsCoreIndex(mg, id2vtx.data(), float(beta));
    } else {
        std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
        return 2;
    }

    if (have_index && !write_bin(out + ".index", index.data(), sizeof(int) * index.size())) return 3;
    FILE* f = std::fopen((out + ".cells").c_str(), "w");
    if (!f) return 3;
    for (const Cell& c : cells) std::fprintf(f, "%u %u %u %.17g %d\n", c.lmd, c.k, c.length, c.density, c.sel);
    std::fclose(f);
    return 0;
}
