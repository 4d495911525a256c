// Caller-side use of the drop-in V3D_CUDA::VocabularyTreeTraversal<BranchFactor, Depth>
// (V3D/CUDA/v3d_cudamatching.h:69-105): allocate, setNodes from a raw tree file (the
// layout VocabularyTree::loadTreeFromFile reads, V3D/Base/v3d_voctree.h:38-51),
// traverseTree on a key file.
// Usage: dropin_voctree tree.bin keys.bin nKeys  -> prints one node id per line.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "v3d_cudamatching.h"

int main(int argc, char** argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s tree.bin keys.bin nKeys\n", argv[0]);
        return 2;
    }
    typedef V3D_CUDA::VocabularyTreeTraversal<10, 3> Traversal;
    int const n = std::atoi(argv[3]);
    std::vector<float> nodes((size_t)Traversal::getNodeCount() * 128), keys((size_t)n * 128);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(nodes.data(), sizeof(float), nodes.size(), f) != nodes.size()) return 1;
    std::fclose(f);
    f = std::fopen(argv[2], "rb");
    if (!f || std::fread(keys.data(), sizeof(float), keys.size(), f) != keys.size()) return 1;
    std::fclose(f);
    try {
        Traversal t;
        t.allocate(n);
        t.setNodes(nodes.data());
        std::vector<int> leaves(n);
        t.traverseTree(n, keys.data(), leaves.data());
        for (int k = 0; k < n; ++k) std::printf("%d\n", leaves[k]);
    } catch (std::exception const& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
