// Batched putative matching of an image sequence: the caller loop of
// V3D/Apps/SfM/match_sequence.cpp:165-240 (every view against the views that follow it
// within 2 * radius - 1, one setLeftFeatures / setRightFeatures / findPutativeMatches per
// pair) turned into batches of pairs for the library's batch API: all images of a window
// of views are uploaded once and every pair of the window is matched in one call.
//
//   v3dm_match_sequence <image list file> <radius> [options]
//
// The image list holds one name per entry; features are read from "<name>.features"
// (V3DBIN or V3DTXT archives, loadFeatures at match_sequence.cpp:25-46).  Per view the
// feature count is the caller's min(maxFeatures, size & 0xffe0) (match_sequence.cpp:183,
// :219), which wraps for 65536 features and more; --exact-count uses size & ~31 instead.
// Matches are written per pair as "view1 view2 count" followed by one "left right" line
// per match, pairs in the caller's order.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "v3dm.h"

namespace {

struct Options {
    std::string list;
    int radius = 0;
    int maxFeatures = 7168;   // MATCHING_MAX_FEATURES default, match_sequence.cpp:134
    float minScore = 0.9f;    // scoreThreshold, match_sequence.cpp:112
    float minRatio = 0.8f;    // MATCHING_SCORE_RATIO default, match_sequence.cpp:123
    int precision = V3DM_PREC_FP32;
    int device = 0;
    int windowViews = 256;    // views whose pairs share one batch
    bool exactCount = false;
    bool quiet = false;
    std::string out;
};

int usage(const char* prog)
{
    fprintf(stderr,
            "Usage: %s <image list file> <radius> [--max-features N] [--min-score S] [--ratio R]\n"
            "          [--precision fp32|bf16] [--device D] [--window V] [--exact-count]\n"
            "          [--out FILE] [--quiet]\n",
            prog);
    return 2;
}

bool parse(int argc, char** argv, Options& o)
{
    if (argc < 3) return false;
    o.list = argv[1];
    o.radius = atoi(argv[2]);
    for (int a = 3; a < argc; ++a) {
        std::string const k = argv[a];
        auto val = [&](const char*& v) {
            if (a + 1 >= argc) return false;
            v = argv[++a];
            return true;
        };
        const char* v = nullptr;
        if (k == "--max-features" && val(v)) o.maxFeatures = atoi(v);
        else if (k == "--min-score" && val(v)) o.minScore = (float)atof(v);
        else if (k == "--ratio" && val(v)) o.minRatio = (float)atof(v);
        else if (k == "--device" && val(v)) o.device = atoi(v);
        else if (k == "--window" && val(v)) o.windowViews = atoi(v);
        else if (k == "--out" && val(v)) o.out = v;
        else if (k == "--precision" && val(v)) {
            std::string const p = v;
            if (p == "fp32") o.precision = V3DM_PREC_FP32;
            else if (p == "bf16") o.precision = V3DM_PREC_BF16;
            else return false;
        } else if (k == "--exact-count") o.exactCount = true;
        else if (k == "--quiet") o.quiet = true;
        else return false;
    }
    return o.radius >= 0 && o.maxFeatures >= 0 && o.windowViews >= 1;
}

struct View {
    int count = -1;             // features used (after the caller's count rule)
    std::vector<float> desc;    // count x 128
};

int fail(const char* what, v3dm_handle h)
{
    fprintf(stderr, "%s failed: %s\n", what, v3dm_last_error(h));
    return 1;
}

}  // namespace

int main(int argc, char** argv)
{
    Options o;
    if (!parse(argc, argv, o)) return usage(argv[0]);

    std::vector<std::string> entries;
    {
        std::ifstream is(o.list);
        if (!is) {
            fprintf(stderr, "cannot read %s\n", o.list.c_str());
            return 1;
        }
        std::string name;
        while (is >> name) entries.push_back(name);
    }
    int const nViews = (int)entries.size();

    v3dm_handle h = nullptr;
    if (v3dm_create(&h, o.device)) return fail("v3dm_create", nullptr);
    if (v3dm_set_precision(h, o.precision)) return fail("v3dm_set_precision", h);
    FILE* out = nullptr;
    if (!o.out.empty()) {
        out = fopen(o.out.c_str(), "w");
        if (!out) {
            fprintf(stderr, "cannot create %s\n", o.out.c_str());
            return 1;
        }
    }

    std::vector<View> views(nViews);
    auto load = [&](int v) -> int {
        if (views[v].count >= 0) return 0;
        std::string const path = entries[v] + ".features";
        if (!o.quiet) printf("reading %s\n", path.c_str());
        int size = 0, binary = 0;
        if (v3dm_features_count(path.c_str(), &size, &binary)) return fail("v3dm_features_count", nullptr);
        int const rule = o.exactCount ? (size & ~31) : (int)(size & 0xffe0);
        int const n = std::min(o.maxFeatures, rule);
        views[v].desc.resize((size_t)n * 128);
        int got = 0;
        if (v3dm_features_read(path.c_str(), n, nullptr, nullptr, nullptr, nullptr, nullptr, views[v].desc.data(),
                               &got))
            return fail("v3dm_features_read", nullptr);
        views[v].count = got;
        return 0;
    };

    long long totalPairs = 0, totalMatches = 0;
    double totalMs = 0.0;
    // windows of views: the pairs (cur, other) with cur in [w0, w1), other in
    // (cur, cur + 2 * radius) -- the loop bounds of match_sequence.cpp:165, :187-189
    for (int w0 = 0; w0 < nViews; w0 += o.windowViews) {
        int const w1 = std::min(nViews, w0 + o.windowViews);
        int const last = std::min(nViews, w1 + 2 * o.radius - 1);  // images the window touches
        std::vector<int32_t> pl, pr;
        for (int cur = w0; cur < w1; ++cur)
            for (int other = cur + 1; other < cur + 2 * o.radius; ++other) {
                if (other >= nViews) continue;
                pl.push_back(cur - w0);
                pr.push_back(other - w0);
            }
        if (pl.empty()) continue;
        std::vector<int32_t> counts;
        std::vector<float> all;
        for (int v = w0; v < last; ++v) {
            if (load(v)) return 1;
            counts.push_back(views[v].count);
            all.insert(all.end(), views[v].desc.begin(), views[v].desc.end());
        }
        if (v3dm_batch_set_images(h, (int)counts.size(), counts.data(), all.data()))
            return fail("v3dm_batch_set_images", h);
        if (v3dm_batch_set_pairs(h, (int)pl.size(), pl.data(), pr.data())) return fail("v3dm_batch_set_pairs", h);
        int cap = 0;
        for (size_t p = 0; p < pl.size(); ++p) cap += std::min(counts[pl[p]], counts[pr[p]]);
        std::vector<int32_t> offsets(pl.size() + 1), pairs(2 * (size_t)std::max(cap, 1));
        int nTotal = 0;
        auto const t0 = std::chrono::steady_clock::now();
        if (v3dm_batch_match(h, o.minScore, o.minRatio, offsets.data(), pairs.data(), std::max(cap, 1), &nTotal))
            return fail("v3dm_batch_match", h);
        totalMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (size_t p = 0; p < pl.size(); ++p) {
            int const a = w0 + pl[p], b = w0 + pr[p];
            int const m = offsets[p + 1] - offsets[p];
            if (!o.quiet) {
                printf("--------------------\n");
                printf("matching view %d with view %d\n", a, b);
                printf("nFeatures1 = %d, nFeatures2 = %d\n", views[a].count, views[b].count);
                printf("matches.size() = %d\n", m);
            }
            if (out) {
                fprintf(out, "%d %d %d\n", a, b, m);
                for (int k = offsets[p]; k < offsets[p + 1]; ++k) fprintf(out, "%d %d\n", pairs[2 * k], pairs[2 * k + 1]);
            }
        }
        totalPairs += (long long)pl.size();
        totalMatches += nTotal;
        // views before the next window are not needed again
        for (int v = w0; v < std::min(w1, nViews); ++v) std::vector<float>().swap(views[v].desc);
    }
    if (out) fclose(out);
    printf("matched %lld pairs (%lld matches) in %.3f ms of batch calls: %.1f pairs/s\n", totalPairs, totalMatches,
           totalMs, totalMs > 0 ? 1e3 * (double)totalPairs / totalMs : 0.0);
    v3dm_destroy(h);
    return 0;
}
