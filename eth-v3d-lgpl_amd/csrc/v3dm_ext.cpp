// Extension hooks: precision modes other than the exact FP32 path.
#include "v3dm_ext.h"

namespace v3dm {

struct ExtState {
    int fallbacks = 0;
};

bool ext_supports(v3dm_context*, int mode) { return mode == V3DM_PREC_FP32; }

int ext_enqueue(v3dm_context* h, int mode, float, float)
{
    (void)mode;
    return ctx_fail(h, V3DM_ERR_UNSUPPORTED, "precision mode not available");
}

int ext_fallbacks(v3dm_context* h)
{
    ExtState* e = ctx_ext(h);
    return e ? e->fallbacks : 0;
}

void ext_features_changed(v3dm_context*, bool) {}

void ext_release(ExtState*) {}

void ext_destroy(ExtState*& e)
{
    delete e;
    e = nullptr;
}

}  // namespace v3dm
