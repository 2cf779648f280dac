// Barnes-Hut on gfx950 (in progress).
#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace orr {

void tree_evaluate(orr_ctx*, const orr_tree_args&, orr_tree_stats*) {
    fail(ORR_ERR_RUNTIME, "tree: not built yet");
}

void tree_copy_nodes(orr_ctx*, double*, double*, double*, double*, double*, std::uint64_t*,
                     std::uint64_t*, std::uint64_t*) {
    fail(ORR_ERR_RUNTIME, "tree: not built yet");
}

}  // namespace orr
