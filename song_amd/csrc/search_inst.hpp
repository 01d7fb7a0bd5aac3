// Instantiation table for search_kernel<METRIC, EXACT, SHAPE>; SHAPE picks the lanes per
// candidate row and the float4 registers per lane from the padded row width.
#pragma once

#include "search_kernel.hpp"

namespace song {

using SearchFn = void (*)(SearchParams);

constexpr int kNumShapes = 5;

// padded row width (floats) -> shape index
inline int shape_for_ld(int ld) {
    if (ld <= 32) return 0;
    if (ld <= 128) return 1;
    if (ld <= 256) return 2;
    if (ld <= 1024) return 3;
    return 4;
}

template <int METRIC, bool EXACT>
inline SearchFn pick_shape(int shape) {
    switch (shape) {
    case 0: return search_kernel<METRIC, EXACT, 0>;
    case 1: return search_kernel<METRIC, EXACT, 1>;
    case 2: return search_kernel<METRIC, EXACT, 2>;
    case 3: return search_kernel<METRIC, EXACT, 3>;
    default: return search_kernel<METRIC, EXACT, 4>;
    }
}

// defined in search_{l2,ip,cos}.hip
SearchFn search_fn_l2(bool exact, int shape);
SearchFn search_fn_ip(bool exact, int shape);
SearchFn search_fn_cos(bool exact, int shape);

}  // namespace song
