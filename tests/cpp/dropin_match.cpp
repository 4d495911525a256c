// Caller-side use of the drop-in V3D_CUDA::SIFT_Matching, written the way the reference's
// match_sequence.cpp drives the matcher (allocate once, set left, set right,
// findPutativeMatches(matches, 0.9f, 0.8f), V3D/Apps/SfM/match_sequence.cpp:136-139,
// :183-188, :227-235).  Synthetic descriptors come from libv3dsynth.
// Usage: dropin_match nLeft nRight pair  -> prints "count" then "left right" lines.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "v3d_cudamatching.h"

extern "C" {
int v3ds_make_pair(uint64_t base, uint32_t config, uint32_t pair, int nL, int nR, double copyFrac, double sigma,
                   uint8_t* qL, uint8_t* qR, int32_t* truth);
void v3ds_dequantise(uint8_t const* q, int n, float* out);
}

struct Vec2 {
    float v[2];
    float operator[](int i) const { return v[i]; }
};

int main(int argc, char** argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s nLeft nRight pair\n", argv[0]);
        return 2;
    }
    int const nL = std::atoi(argv[1]), nR = std::atoi(argv[2]), pair = std::atoi(argv[3]);
    std::vector<uint8_t> qL((size_t)nL * 128), qR((size_t)nR * 128);
    std::vector<int32_t> truth(nR);
    v3ds_make_pair(0x5EEDE7A3DULL, 0, (uint32_t)pair, nL, nR, 0.4, 0.05, qL.data(), qR.data(), truth.data());
    std::vector<float> L((size_t)nL * 128), R((size_t)nR * 128);
    v3ds_dequantise(qL.data(), nL, L.data());
    v3ds_dequantise(qR.data(), nR, R.data());

    try {
        V3D_CUDA::SIFT_Matching matcher;
        matcher.enableRoundingWarning(false);
        matcher.allocate(7168, false);
        int const n1 = std::min(7168, nL & 0xffe0);
        int const n2 = std::min(7168, nR & 0xffe0);
        matcher.setLeftFeatures(n1, L.data());
        matcher.setRightFeatures(n2, R.data());
        std::vector<std::pair<int, int> > matches;
        matcher.findPutativeMatches(matches, 0.9f, 0.8f);
        std::printf("%zu\n", matches.size());
        for (auto const& m : matches) std::printf("%d %d\n", m.first, m.second);
        matcher.deallocate();
    } catch (std::exception const& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
