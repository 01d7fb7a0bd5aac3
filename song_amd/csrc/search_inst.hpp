// Instantiation table for search_kernel / search_persistent<METRIC, EXACT, STORE, SHAPE>.  SHAPE picks the
// lanes per candidate row and the 16-byte units per lane from the units per row; EXACT
// (double accumulation, bit-exact to the CPU) exists for float32 rows only.
#pragma once

#include "search_kernel.hpp"

namespace song {

using SearchFn = void (*)(SearchParams);

// 16-byte units per row -> shape index
inline int shape_for_units(int units) {
    if (units <= 8) return 0;
    if (units <= 32) return 1;
    if (units <= 64) return 2;
    if (units <= 256) return 3;
    return 4;
}

// bagl: the persistent form with the queue's bag in LDS; bg: the best list's front in global
// memory (both float32 rows only; null otherwise)
template <int METRIC, bool EXACT, int STORE, int SHAPE>
inline SearchFn pick_form(bool persistent, bool bagl, bool bg) {
    if (bagl && bg) return nullptr;
    if (bagl) {
        if constexpr (STORE == ST_F32) return persistent ? search_persistent<METRIC, EXACT, STORE, SHAPE, true> : nullptr;
        return nullptr;
    }
    if (bg) {
        if constexpr (STORE == ST_F32)
            return persistent ? search_persistent<METRIC, EXACT, STORE, SHAPE, false, true>
                              : search_kernel<METRIC, EXACT, STORE, SHAPE, true>;
        return nullptr;
    }
    return persistent ? search_persistent<METRIC, EXACT, STORE, SHAPE> : search_kernel<METRIC, EXACT, STORE, SHAPE>;
}

template <int METRIC, bool EXACT, int STORE>
inline SearchFn pick_shape(int shape, bool persistent, bool bagl, bool bg) {
    switch (shape) {
    case 0: return pick_form<METRIC, EXACT, STORE, 0>(persistent, bagl, bg);
    case 1: return pick_form<METRIC, EXACT, STORE, 1>(persistent, bagl, bg);
    case 2: return pick_form<METRIC, EXACT, STORE, 2>(persistent, bagl, bg);
    case 3: return pick_form<METRIC, EXACT, STORE, 3>(persistent, bagl, bg);
    default: return pick_form<METRIC, EXACT, STORE, 4>(persistent, bagl, bg);
    }
}

// two waves per query (search_duo): float32 keys on float32 rows
template <int METRIC, bool BG>
inline SearchFn pick_duo_bg(int shape) {
    switch (shape) {
    case 0: return search_duo<METRIC, ST_F32, 0, BG>;
    case 1: return search_duo<METRIC, ST_F32, 1, BG>;
    case 2: return search_duo<METRIC, ST_F32, 2, BG>;
    case 3: return search_duo<METRIC, ST_F32, 3, BG>;
    default: return search_duo<METRIC, ST_F32, 4, BG>;
    }
}
template <int METRIC>
inline SearchFn pick_duo(int shape, bool bg) {
    return bg ? pick_duo_bg<METRIC, true>(shape) : pick_duo_bg<METRIC, false>(shape);
}

template <int METRIC>
inline SearchFn pick_kernel(bool exact, int store, int shape, bool persistent, bool bagl, bool bg) {
    if (exact) return pick_shape<METRIC, true, ST_F32>(shape, persistent, bagl, bg);
    switch (store) {
    case ST_BF16: return pick_shape<METRIC, false, ST_BF16>(shape, persistent, bagl, bg);
    case ST_FP8: return pick_shape<METRIC, false, ST_FP8>(shape, persistent, bagl, bg);
    default: return pick_shape<METRIC, false, ST_F32>(shape, persistent, bagl, bg);
    }
}

// defined in search_{l2,ip,cos}.hip
// persistent: the work-list kernel search_persistent instead of one group per query
// bagl: the queue's bag in LDS (persistent form, float32 rows)
// bg: the best list's front in global memory (float32 rows)
SearchFn search_fn_l2(bool exact, int store, int shape, bool persistent, bool bagl, bool bg);
SearchFn search_fn_ip(bool exact, int store, int shape, bool persistent, bool bagl, bool bg);
SearchFn search_fn_cos(bool exact, int store, int shape, bool persistent, bool bagl, bool bg);
SearchFn duo_fn_l2(int shape, bool bg);
SearchFn duo_fn_ip(int shape, bool bg);
SearchFn duo_fn_cos(int shape, bool bg);
// defined in group_{l2,ip,cos}.hip: lanes = 16 or 32 lanes per query
SearchFn group_fn_l2(bool exact, int store, int shape, int lanes);
SearchFn group_fn_ip(bool exact, int store, int shape, int lanes);
SearchFn group_fn_cos(bool exact, int store, int shape, int lanes);

}  // namespace song
